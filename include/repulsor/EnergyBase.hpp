#pragma once
// Polymorphic energy interface of the public API.
//
// EnergyBase<SimplicialMeshBase<...>> (reference src/Energies/EnergyBase.hpp:6-176): Value(M) and
// Differential(M) with the mesh's value cache, the two uncached pointer forms of Differential, and the
// protected virtual value/differential that a concrete energy implements.
// EnergyDimRestricted<Mesh_T> (reference src/Energies/EnergyDimRestricted.hpp:6-124): downcasts the mesh
// to the concrete SimplicialMesh type and forwards; a mesh of another type gets an error message and 0.
//
// Errors never throw across this API: a failing computation prints a message and returns 0, leaving the
// output untouched, as the reference does.

#include "SimplicialMeshBase.hpp"

namespace Repulsor
{
    template<typename MeshBase_T> class EnergyBase {};

    template<typename Real, typename Int, typename LInt, typename SReal, typename ExtReal>
    class EnergyBase<SimplicialMeshBase<Real,Int,LInt,SReal,ExtReal>>
    {
    public:
        using MeshBase_T        = SimplicialMeshBase<Real,Int,LInt,SReal,ExtReal>;
        using TangentVector_T   = typename MeshBase_T::TangentVector_T;
        using CotangentVector_T = typename MeshBase_T::CotangentVector_T;

        EnergyBase() = default;
        virtual ~EnergyBase() = default;

    protected:
        // Whether the last value / differential call of this object succeeded.
        mutable bool last_ok = true;

    public:
        // Energy value, cached in the mesh until its next SemiStaticUpdate.
        ExtReal Value( cref<MeshBase_T> M ) const
        {
            const std::string tag = ClassName() + "::Value";
            ScopedTimer timer( tag );
            auto it = M.scalar_cache.find( tag );
            if( it != M.scalar_cache.end() ) { return it->second; }
            last_ok = true;
            const ExtReal en = value( M );
            if( last_ok ) { M.scalar_cache[tag] = en; }
            return en;
        }

        // Vertex differential (VertexCount x AmbDim), cached in the mesh until its next SemiStaticUpdate;
        // the value computed with it is cached too.
        cref<CotangentVector_T> Differential( cref<MeshBase_T> M ) const
        {
            const std::string tag = ClassName() + "::Differential";
            ScopedTimer timer( tag );
            auto it = M.tensor_cache.find( tag );
            if( it != M.tensor_cache.end() ) { return it->second; }
            last_ok = true;
            const ExtReal en = differential( M );
            CotangentVector_T diff( M.VertexCount(), M.AmbDim(), ExtReal(0) );
            if( !last_ok ) { return empty; }
            M.Assemble_ClusterTree_Derivatives( ExtReal(1), ExtReal(0), diff.data(), M.AmbDim() );
            M.scalar_cache.emplace( ClassName() + "::Value", en );
            return M.tensor_cache.emplace( tag, std::move( diff ) ).first->second;
        }

        // Y = alpha * dE + beta * Y (row-major, leading dimension ldY); returns the energy. No caching.
        ExtReal Differential( cref<MeshBase_T> M, const ExtReal alpha, const ExtReal beta, mptr<ExtReal> Y, const Int ldY ) const
        {
            ScopedTimer timer( ClassName() + "::Differential (pointer)" );
            last_ok = true;
            const ExtReal en = differential( M );
            if( !last_ok ) { return ExtReal(0); }
            M.Assemble_ClusterTree_Derivatives( alpha, beta, Y, ldY );
            return en;
        }

        // Y = dE (+ Y with addto), leading dimension AmbDim; returns the energy. No caching.
        ExtReal Differential( cref<MeshBase_T> M, mptr<ExtReal> Y, const bool addto = false ) const
        {
            return Differential( M, ExtReal(1), addto ? ExtReal(1) : ExtReal(0), Y, M.AmbDim() );
        }

        static std::string className() { return "EnergyBase<double,int32,LInt,double,double>"; }
        virtual std::string ClassName() const { return className(); }

    protected:
        const CotangentVector_T empty;

        // Energy of M; on failure print a message, set last_ok = false and return 0.
        virtual ExtReal value( cref<MeshBase_T> M ) const = 0;
        // Energy of M and its differential, held for M.Assemble_ClusterTree_Derivatives; on failure print
        // a message, set last_ok = false and return 0.
        virtual ExtReal differential( cref<MeshBase_T> M ) const = 0;
    };

    template<typename Mesh_T>
    class EnergyDimRestricted : public EnergyBase<typename Mesh_T::Base_T>
    {
    public:
        using MeshBase_T = typename Mesh_T::Base_T;
        using Base_T     = EnergyBase<MeshBase_T>;
        using ExtReal    = typename MeshBase_T::ExtReal;

        using Base_T::Value;
        using Base_T::Differential;

        EnergyDimRestricted() = default;
        ~EnergyDimRestricted() override = default;

        static std::string className() { return "EnergyDimRestricted<" + std::to_string( Mesh_T::DOM_DIM ) + "," + std::to_string( Mesh_T::AMB_DIM ) + ">"; }
        std::string ClassName() const override { return className(); }

    protected:
        using Base_T::last_ok;

        ExtReal value( cref<MeshBase_T> M ) const override
        {
            const Mesh_T * Q = dynamic_cast<const Mesh_T *>( &M );
            if( Q == nullptr )
            {
                eprint( this->ClassName() + "::value: Input could not be downcast to compatible type. Doing nothing." );
                last_ok = false;
                return ExtReal(0);
            }
            return value( *Q );
        }

        ExtReal differential( cref<MeshBase_T> M ) const override
        {
            const Mesh_T * Q = dynamic_cast<const Mesh_T *>( &M );
            if( Q == nullptr )
            {
                eprint( this->ClassName() + "::differential: Input could not be downcast to compatible type. Doing nothing." );
                last_ok = false;
                return ExtReal(0);
            }
            return differential( *Q );
        }

        virtual ExtReal value( cref<Mesh_T> M ) const = 0;
        virtual ExtReal differential( cref<Mesh_T> M ) const = 0;
    };

} // namespace Repulsor
