#pragma once
// Polymorphic mesh interface of the public API, with the reference's template parameters and the members
// that its drivers and the energy/metric base classes call (reference
// src/SimplicialMesh/SimplicialMeshBase.hpp:6-207): dimensions and counts, vertex coordinates and
// simplices, the settings members, SemiStaticUpdate, Assemble_ClusterTree_Derivatives,
// MaximumSafeStepSize, the H^1 weights and solver ID, and the value cache that EnergyBase uses
// (the reference's CachedObject: InCacheQ / cached values cleared by SemiStaticUpdate).
//
// Supported types: Real = SReal = ExtReal = double (the device path computes in FP64, or in FP32 with
// FP64 accumulation as chosen per energy/metric object) and a 32-bit signed Int (device indices are
// int32); LInt is any integer type (row pointers are int64 on the device).

#include "Common.hpp"
#include "Settings.hpp"
#include "Tensors.hpp"

#include <map>
#include <type_traits>

namespace Repulsor
{
    template<typename Real_, typename Int_, typename LInt_, typename SReal_ = Real_, typename ExtReal_ = Real_>
    class SimplicialMeshBase
    {
        static_assert( std::is_same<Real_, double>::value && std::is_same<SReal_, double>::value && std::is_same<ExtReal_, double>::value,
                       "SimplicialMeshBase: Real, SReal and ExtReal must be double" );
        static_assert( std::is_integral<Int_>::value && std::is_signed<Int_>::value && sizeof(Int_) == 4,
                       "SimplicialMeshBase: Int must be a 32-bit signed integer type" );
        static_assert( std::is_integral<LInt_>::value, "SimplicialMeshBase: LInt must be an integer type" );

    public:
        using Real    = Real_;
        using Int     = Int_;
        using LInt    = LInt_;
        using SReal   = SReal_;
        using ExtReal = ExtReal_;

        using TangentVector_T   = Tensor2<ExtReal,Int>;
        using CotangentVector_T = Tensor2<ExtReal,Int>;

        SimplicialMeshBase() = default;
        explicit SimplicialMeshBase( const Int thread_count_ ) : thread_count( std::max( Int(1), thread_count_ ) ) {}
        virtual ~SimplicialMeshBase() = default;

        SimplicialMeshBase( const SimplicialMeshBase & ) = delete;
        SimplicialMeshBase & operator=( const SimplicialMeshBase & ) = delete;

        mutable ClusterTreeSettings      cluster_tree_settings;
        mutable BlockClusterTreeSettings block_cluster_tree_settings;
        mutable AdaptivitySettings       adaptivity_settings;

    protected:
        Int thread_count = 1;

    public:
        virtual Int DomDim() const = 0;
        virtual Int AmbDim() const = 0;
        virtual cref<Tensor2<Real,Int>> VertexCoordinates() const = 0;
        virtual cref<Tensor2<Int,Int>> Simplices() const = 0;
        virtual Int FarDim() const = 0;
        virtual Int NearDim() const = 0;
        virtual Int VertexCount() const = 0;
        virtual Int SimplexCount() const = 0;
        virtual Int ThreadCount() const { return thread_count; }

        // New vertex coordinates (V x AmbDim row-major, or AmbDim x V with transp) without rebuilding the
        // cluster tree and the block cluster tree; clears the cached values.
        virtual void SemiStaticUpdate( cptr<ExtReal> V_coords_, const bool transp_ = false ) const = 0;

        // Y = alpha * (vertex differential of the last Differential computation) + beta * Y.
        virtual void Assemble_ClusterTree_Derivatives( const ExtReal alpha, const ExtReal beta, mptr<ExtReal> Y, const Int ldY ) const = 0;

        // Largest t in [0, max_time] for which moving every vertex by t vecs makes no two simplices without
        // a common vertex meet (bisection to relative tolerance TOL).
        virtual ExtReal MaximumSafeStepSize( cptr<ExtReal> vecs, const ExtReal max_time, const ExtReal TOL = ExtReal(0.0625),
                                             const bool transp_ = false ) = 0;

        virtual Real H1StiffnessWeight() const = 0;
        virtual void SetH1StiffnessWeight( const Real weight ) = 0;
        virtual Real H1MassWeight() const = 0;
        virtual void SetH1MassWeight( const Real weight ) = 0;

        // Fill-reducing ordering of the H^1 Cholesky factorisation (0: nested dissection on the cluster
        // tree, the reference's default).
        virtual Int H1SolverID() const = 0;
        virtual void SetH1SolverID( const Int ID ) const = 0;

        // ---------------------------------------------------------------- value cache
        // Values keyed by ClassName() + "::Value" / "::Differential", cleared by SemiStaticUpdate.
        mutable std::map<std::string,double> scalar_cache;
        mutable std::map<std::string,CotangentVector_T> tensor_cache;

        bool InCacheQ( const std::string & tag ) const { return scalar_cache.count( tag ) || tensor_cache.count( tag ); }

        virtual void ClearCache() const
        {
            scalar_cache.clear();
            tensor_cache.clear();
        }

        static std::string className() { return "SimplicialMeshBase<double,int32,LInt,double,double>"; }
        virtual std::string ClassName() const { return className(); }
    };

} // namespace Repulsor
