#pragma once
// Simplicial mesh of dimension DOM_DIM in R^3 with the reference's caching behaviour
// (src/SimplicialMesh.hpp:13-869): the cluster tree and the block cluster tree are persistent
// (built once from the coordinates frozen at construction), while energy/metric values are cleared
// by SemiStaticUpdate. The GPU context (librepulsor_hip.so) is owned by the mesh, like the
// reference's persistent cache entries.

#include "Common.hpp"
#include "Settings.hpp"
#include "SimplicialMeshBase.hpp"
#include "SimplexData.hpp"
#include "ClusterTree.hpp"
#include "BlockClusterTree.hpp"
#include "H1Metric.hpp"
#include "SparseCholesky.hpp"
#include "../repulsor_hip.h"

#include <map>
#include <unordered_map>

namespace Repulsor
{
    // RAII wrapper of an rp_ctx.
    class DeviceContext
    {
        rp_ctx * ctx = nullptr;
    public:
        explicit DeviceContext( int device )
        {
            if( rp_ctx_create( device, &ctx ) != RP_OK )
            {
                ctx = nullptr;
                throw std::runtime_error( std::string("rp_ctx_create failed: ") + rp_last_error() );
            }
        }
        ~DeviceContext() { if( ctx ) { rp_ctx_destroy( ctx ); } }
        DeviceContext( const DeviceContext & ) = delete;
        DeviceContext & operator=( const DeviceContext & ) = delete;
        rp_ctx * get() const { return ctx; }
    };

    inline bool Check( int status, const std::string & where )
    {
        if( status != RP_OK ) { eprint( where + ": " + rp_last_error() ); return false; }
        return true;
    }

    // SimplicialMesh<DOM_DIM,AMB_DIM,Real,Int,LInt,SReal,ExtReal> (reference src/SimplicialMesh.hpp:13-869)
    // for curves (DOM_DIM 1) and surfaces (DOM_DIM 2) in R^3. The template parameters after AMB_DIM are the
    // interface types of SimplicialMeshBase (defaults: double, int32, int64); internal arrays use int32
    // indices and int64 offsets.
    template<int DOM_DIM_, int AMB_DIM_ = 3, typename Real_ = double, typename Int_ = int32_t, typename LInt_ = int64_t,
             typename SReal_ = Real_, typename ExtReal_ = Real_>
    class SimplicialMesh : public SimplicialMeshBase<Real_,Int_,LInt_,SReal_,ExtReal_>
    {
        static_assert( AMB_DIM_ == 3, "only meshes in R^3 are supported" );
        static_assert( DOM_DIM_ == 1 || DOM_DIM_ == 2, "only curves and surfaces are supported" );

    public:
        static constexpr int DOM_DIM  = DOM_DIM_;
        static constexpr int AMB_DIM  = AMB_DIM_;
        static constexpr int SIZE     = DOM_DIM + 1;
        static constexpr int NEAR_DIM = SimplexData<DOM_DIM>::NEAR_DIM;
        static constexpr int FAR_DIM  = SimplexData<DOM_DIM>::FAR_DIM;

        using Base_T            = SimplicialMeshBase<Real_,Int_,LInt_,SReal_,ExtReal_>;
        using ExtInt_T          = Int_;
        using TangentVector_T   = typename Base_T::TangentVector_T;
        using CotangentVector_T = typename Base_T::CotangentVector_T;

        using Tree_T = ClusterTree<DOM_DIM>;
        using BCT_T  = BlockClusterTree<DOM_DIM>;

        using Base_T::cluster_tree_settings;
        using Base_T::block_cluster_tree_settings;
        using Base_T::adaptivity_settings;
        using Base_T::scalar_cache;
        using Base_T::tensor_cache;

    protected:
        int threads = 1;
        int device  = 0;
        mutable Tensor2<double,Int_> V_coords;        // host copy (refreshed from the device when stale)
        std::vector<double> V_coords_frozen, V_charges;
        Tensor2<Int_,Int_>  simplices;
        bool have_charges = false;

        mutable std::unique_ptr<Tree_T> tree;
        mutable std::unique_ptr<BCT_T>  bct;

        // device state
        mutable std::unique_ptr<DeviceContext> dev;
        mutable bool dev_topology = false;
        mutable bool dev_data     = false;
        mutable bool dev_h1_structure = false;   // the device holds the structure of h1_factor (its values may be older)
        // after a device-resident SemiStaticUpdate the device holds newer coordinates (coords_stale);
        // after any SemiStaticUpdate the host tree rows are recomputed only when a host computation
        // needs them (rows_stale)
        mutable bool coords_stale = false;
        mutable bool rows_stale   = false;

    public:
        // Row-major V x 3 coordinates and T x (n+1) simplices, optional vertex charges.
        template<typename ExtReal, typename ExtInt>
        SimplicialMesh( const ExtReal * V_coords_, const LInt vertex_count_,
                        const ExtInt * simplices_, const LInt simplex_count_,
                        const LInt thread_count_ = 1, const double * vertex_charges_ = nullptr )
        :   Base_T( (Int_)std::max<LInt>( 1, thread_count_ ) ), threads( (int)std::max<LInt>( 1, thread_count_ ) )
        {
            Read( V_coords_, vertex_count_, false, simplices_, simplex_count_, false );
            if( vertex_charges_ )
            {
                have_charges = true;
                V_charges.assign( vertex_charges_, vertex_charges_ + vertex_count_ );
            }
        }

        // The reference's constructor (src/SimplicialMesh.hpp:143-185, called by SimplicialMesh_Factory::Make):
        // coordinates V x AMB_DIM and simplices T x (DOM_DIM+1), each row-major or, with the ColMajor flag,
        // column-major (AMB_DIM x V, (DOM_DIM+1) x T).
        template<typename ExtReal, typename ExtInt>
        SimplicialMesh( const ExtReal * V_coords_, const Int_ vertex_count_, const bool V_ColMajorQ,
                        const ExtInt * simplices_, const Int_ simplex_count_, const bool S_ColMajorQ,
                        const Int_ thread_count_ = 1 )
        :   Base_T( std::max<Int_>( 1, thread_count_ ) ), threads( (int)std::max<Int_>( 1, thread_count_ ) )
        {
            Read( V_coords_, vertex_count_, V_ColMajorQ, simplices_, simplex_count_, S_ColMajorQ );
        }

    private:
        template<typename ExtReal, typename ExtInt>
        void Read( const ExtReal * V_, const LInt V, const bool V_transp, const ExtInt * S_, const LInt T, const bool S_transp )
        {
            if( V <= 0 || T <= 0 || !V_ || !S_ ) { throw std::invalid_argument( "SimplicialMesh: empty mesh" ); }
            V_coords = Tensor2<double,Int_>( (Int_)V, (Int_)AMB_DIM );
            for( LInt v = 0; v < V; ++v )
            {
                for( int k = 0; k < AMB_DIM; ++k ) { V_coords( (Int_)v, k ) = (double)( V_transp ? V_[k * V + v] : V_[AMB_DIM * v + k] ); }
            }
            V_coords_frozen.assign( V_coords.begin(), V_coords.end() );
            simplices = Tensor2<Int_,Int_>( (Int_)T, (Int_)SIZE );
            for( LInt t = 0; t < T; ++t )
            {
                for( int k = 0; k < SIZE; ++k )
                {
                    const LInt idx = (LInt)( S_transp ? S_[k * T + t] : S_[SIZE * t + k] );
                    if( idx < 0 || idx >= V ) { throw std::invalid_argument( "SimplicialMesh: simplex vertex index out of range" ); }
                    simplices( (Int_)t, k ) = (Int_)idx;
                }
            }
            cluster_tree_settings.thread_count = threads;
        }

    public:
        Int_ VertexCount()  const override { return V_coords.Dim(0); }
        Int_ SimplexCount() const override { return simplices.Dim(0); }
        Int_ AmbDim()       const override { return AMB_DIM; }
        Int_ DomDim()       const override { return DOM_DIM; }
        Int_ FarDim()       const override { return FAR_DIM; }
        Int_ NearDim()      const override { return NEAR_DIM; }
        Int_ ThreadCount()  const override { return threads; }
        void SetDevice( int d ) { device = d; }

        cref<Tensor2<double,Int_>> VertexCoordinates() const override { RequireHostCoordinates(); return V_coords; }
        cref<Tensor2<Int_,Int_>> Simplices() const override { return simplices; }

        static std::string className()
        {
            return "SimplicialMesh<" + std::to_string( DOM_DIM ) + "," + std::to_string( AMB_DIM ) + ",double,int32,LInt,double,double>";
        }
        std::string ClassName() const override { return className(); }

        // Host copy of the coordinates of the current state.
        void RequireHostCoordinates() const
        {
            if( !coords_stale ) { return; }
            if( rp_download_state_rows( dev->get(), nullptr, nullptr, V_coords.data() ) != RP_OK )
            {
                throw std::runtime_error( std::string("rp_download_state_rows: ") + rp_last_error() );
            }
            coords_stale = false;
        }
        const double * VertexCharges() const { return have_charges ? V_charges.data() : nullptr; }

        // ---------------------------------------------------------------- persistent cache
        const Tree_T & GetClusterTree() const
        {
            RequireHostCoordinates();
            if( tree && rows_stale )
            {
                ScopedTimer timer("ClusterTree::SemiStaticUpdate");
                const LInt T = SimplexCount();
                std::vector<double> P_near( (size_t)(NEAR_DIM * T) ), P_far( (size_t)(FAR_DIM * T) );
                ComputeNearFarData<DOM_DIM>( V_coords.data(), simplices.data(), VertexCharges(), T, P_near.data(), P_far.data(), threads );
                tree->SemiStaticUpdate( P_near.data(), P_far.data() );
            }
            rows_stale = false;
            if( !tree )
            {
                ScopedTimer timer("SimplicialMesh::GetClusterTree");
                std::unique_ptr<BuiltClusterTree> built;
                if( cluster_tree_settings.device_build ) { built = BuildTopologyOnDevice(); }
                const LInt T = SimplexCount();
                std::vector<double> P_near( (size_t)(NEAR_DIM * T) ), P_far( (size_t)(FAR_DIM * T) );
                std::vector<double> diffop( (size_t)(3 * SIZE * T) );
                std::vector<Int> sorted( (size_t)(SIZE * T) );
                ParallelFor( T, threads, [&]( LInt i )
                {
                    SimplexData<DOM_DIM> ker;
                    ker.Read( V_coords.data(), simplices.data(), VertexCharges(), i );
                    ker.WriteNear( &P_near[(size_t)(NEAR_DIM * i)] );
                    ker.WriteFar ( &P_far [(size_t)(FAR_DIM  * i)] );
                    ker.WriteDiffOp( &diffop[(size_t)(3 * SIZE * i)] );
                    for( int l = 0; l < SIZE; ++l ) { sorted[(size_t)(SIZE * i + l)] = ker.s_simplex[l]; }
                });
                ClusterTreeSettings s = cluster_tree_settings;
                if( s.thread_count <= 0 ) { s.thread_count = threads; }
                tree = std::make_unique<Tree_T>( V_coords_frozen.data(), simplices.data(), T,
                                                 P_near.data(), P_far.data(), diffop.data(), sorted.data(),
                                                 s, threads, built.get() );
            }
            return *tree;
        }

        const BCT_T & GetBlockClusterTree() const
        {
            if( !bct )
            {
                ScopedTimer timer("SimplicialMesh::GetBlockClusterTree");
                // The adaptivity settings overwrite the near-field parameters (src/SimplicialMesh.hpp:592-595).
                BlockClusterTreeSettings s = block_cluster_tree_settings;
                s.near_field_separation_parameter   = adaptivity_settings.theta;
                s.near_field_intersection_parameter = adaptivity_settings.intersection_theta;
                s.max_refinement                    = adaptivity_settings.max_refinement;
                const Tree_T & S = GetClusterTree();
                if( built_bct ) { bct = std::make_unique<BCT_T>( S, s, threads, std::move( *built_bct ) ); built_bct.reset(); }
                else            { bct = std::make_unique<BCT_T>( S, s, threads ); }
            }
            return *bct;
        }

    protected:
        mutable std::unique_ptr<BuiltBlockClusterTree> built_bct;   // lists of the device build until the BCT adopts them

        // Cluster tree and block cluster tree on the GPU from the frozen coordinates (rp_build_topology),
        // with the block cluster tree settings in force now (the near parameters from the adaptivity
        // settings, as GetBlockClusterTree takes them).
        std::unique_ptr<BuiltClusterTree> BuildTopologyOnDevice() const
        {
            rp_ctx * c = Context();
            const LInt T = SimplexCount();
            int64_t cnt[8] = {};
            const double near_theta = adaptivity_settings.theta;
            if( rp_build_topology( c, DOM_DIM, VertexCount(), T, V_coords_frozen.data(), simplices.data(), cluster_tree_settings.split_threshold,
                                   block_cluster_tree_settings.far_field_separation_parameter, near_theta, near_theta, cnt ) != RP_OK )
            {
                throw std::runtime_error( std::string("rp_build_topology: ") + rp_last_error() );
            }
            const size_t C = (size_t)cnt[0], L = (size_t)cnt[1];
            auto B = std::make_unique<BuiltClusterTree>();
            B->P_ordering.resize( (size_t)T );
            B->C_left.resize( C ); B->C_right.resize( C ); B->C_begin.resize( C ); B->C_end.resize( C ); B->C_depth.resize( C );
            B->C_serialized.resize( 7 * C );
            built_bct = std::make_unique<BuiltBlockClusterTree>();
            BuiltBlockClusterTree & F = *built_bct;
            F.far_rp.resize( C + 1 ); F.far_ci.resize( (size_t)cnt[3] );
            std::vector<Int> blk( 2 * (size_t)cnt[4] ), vf( 2 * (size_t)cnt[5] ), inter( 2 * (size_t)cnt[6] );
            F.col_ptr.resize( L + 1 ); F.col_prim.resize( (size_t)cnt[7] ); F.col_mask.resize( (size_t)cnt[7] );
            if( rp_download_built_topology( c, B->P_ordering.data(), B->C_left.data(), B->C_right.data(), B->C_begin.data(), B->C_end.data(),
                                            B->C_depth.data(), B->C_serialized.data(), F.far_rp.data(), F.far_ci.data(), blk.data(), vf.data(),
                                            inter.data(), F.col_ptr.data(), F.col_prim.data(), F.col_mask.data(), 1 ) != RP_OK )
            {
                built_bct.reset();
                throw std::runtime_error( std::string("rp_download_built_topology: ") + rp_last_error() );
            }
            auto to_pairs = []( const std::vector<Int> & a, std::vector<std::pair<Int,Int>> & out )
            {
                out.resize( a.size() / 2 );
                for( size_t k = 0; k < out.size(); ++k ) { out[k] = { a[2 * k], a[2 * k + 1] }; }
            };
            to_pairs( blk, F.leaf_blocks ); to_pairs( vf, F.verynear_pairs ); to_pairs( inter, F.inter_pairs );
            return B;
        }

    public:
        // Reads new coordinates without rebuilding the tree (src/SimplicialMesh.hpp:354-376).
        // With device_ptr, V_new is a V x 3 row-major array in HBM and the state is computed on the GPU
        // (rp_update_coordinates); the host reads it back only when a host computation needs it
        // (H^1 metric, learned-preconditioner features, read-outs). Without it, the coordinates are
        // copied and, once the device is set up, sent to the device, which computes the per-state data
        // itself (15.7 MB of coordinates instead of 168 MB of rows at 1.31M triangles).
        // The reference's form: errors are printed, not thrown.
        void SemiStaticUpdate( const double * V_new, const bool transp = false ) const override
        {
            try { SemiStaticUpdate( V_new, transp, false ); }
            catch( const std::exception & e ) { eprint( ClassName() + "::SemiStaticUpdate: " + e.what() ); }
        }

        void SemiStaticUpdate( const double * V_new, const bool transp, const bool device_ptr ) const
        {
            ScopedTimer timer("SimplicialMesh::SemiStaticUpdate");
            const LInt V = VertexCount();
            ClearCache();
            if( device_ptr )
            {
                if( transp ) { throw std::runtime_error( "SemiStaticUpdate: device coordinates must be row-major (transp = false)" ); }
                rp_ctx * c = Device();
                const int singletons = GetClusterTree().SplitThreshold() <= 1 ? 1 : 0;
                if( rp_update_coordinates( c, V_new, 3, singletons, RP_DEVICE_PTR ) != RP_OK )
                {
                    throw std::runtime_error( std::string("rp_update_coordinates: ") + rp_last_error() );
                }
                coords_stale = true;
                rows_stale = (bool)tree;
                dev_data = true;
                return;
            }
            if( transp ) { for( LInt v = 0; v < V; ++v ) { for( int k = 0; k < 3; ++k ) { V_coords( (Int_)v, k ) = V_new[k * V + v]; } } }
            else { std::copy( V_new, V_new + 3 * V, V_coords.begin() ); }
            coords_stale = false;
            rows_stale = (bool)tree;
            if( dev && dev_topology )
            {
                const int singletons = GetClusterTree().SplitThreshold() <= 1 ? 1 : 0;
                if( rp_update_coordinates( dev->get(), V_coords.data(), 3, singletons, 0 ) != RP_OK )
                {
                    throw std::runtime_error( std::string("rp_update_coordinates: ") + rp_last_error() );
                }
                dev_data = true;
                device_state.clear();
            }
            else { dev_data = false; }
        }

        // Largest t in [0, max_time] such that moving every vertex by t vecs causes no collision between
        // simplices without a common vertex (SimplicialMesh::MaximumSafeStepSize, src/SimplicialMesh.hpp:441-468).
        // vecs: V x 3 (3 x V with transp), host memory, or V x 3 row-major in HBM with device_ptr.
        // The reference's form: errors are printed, not thrown, and give 0.
        double MaximumSafeStepSize( const double * vecs, const double max_time, const double TOL = 0.0625,
                                    const bool transp = false ) override
        {
            try { return static_cast<const SimplicialMesh &>( *this ).MaximumSafeStepSize( vecs, max_time, TOL, transp, false ); }
            catch( const std::exception & e ) { eprint( ClassName() + "::MaximumSafeStepSize: " + e.what() ); return 0.0; }
        }

        double MaximumSafeStepSize( const double * vecs, const double max_time, const double TOL,
                                    const bool transp, const bool device_ptr ) const
        {
            ScopedTimer timer("SimplicialMesh::MaximumSafeStepSize");
            rp_ctx * c = Device();
            std::vector<double> buf;
            const double * U = vecs;
            if( transp )
            {
                if( device_ptr ) { throw std::runtime_error( "MaximumSafeStepSize: device update vectors must be row-major" ); }
                const LInt V = VertexCount();
                buf.resize( (size_t)(3 * V) );
                for( LInt v = 0; v < V; ++v ) { for( int k = 0; k < 3; ++k ) { buf[(size_t)(3 * v + k)] = vecs[k * V + v]; } }
                U = buf.data();
            }
            double t = 0.0;
            if( rp_max_safe_step( c, U, 3, max_time, TOL, device_ptr ? RP_DEVICE_PTR : 0, &t, last_collision_stats ) != RP_OK )
            {
                throw std::runtime_error( std::string("rp_max_safe_step: ") + rp_last_error() );
            }
            return t;
        }

        // Leaf pairs, simplex pairs tested and hull tests of the last MaximumSafeStepSize.
        const int64_t * CollisionStats() const { return last_collision_stats; }

    protected:
        mutable int64_t last_collision_stats[3] = { 0, 0, 0 };

    public:
        // ---------------------------------------------------------------- value cache
        mutable std::map<std::string,int> device_state;   // e.g. which metric the device holds

        void ClearCache() const override
        {
            Base_T::ClearCache();
            device_state.clear();
            h1.reset();
            // a moved mesh keeps the H^1 pattern: the factor keeps its symbolic structure and is refactorised
            // numerically on its next use
            if( h1_factor ) { h1_values_stale = true; }
        }

        // Y = alpha * dE + beta * Y for the differential last computed on the device
        // (SimplicialMesh::Assemble_ClusterTree_Derivatives, src/SimplicialMesh.hpp:651-675).
        void Assemble_ClusterTree_Derivatives( const double alpha, const double beta, double * Y, const Int_ ldY ) const override
        {
            if( !dev ) { eprint( ClassName() + "::Assemble_ClusterTree_Derivatives: no differential computed. Doing nothing." ); return; }
            Check( rp_assemble_vertex_gradient( dev->get(), alpha, beta, Y, ldY, 0 ), ClassName() + "::Assemble_ClusterTree_Derivatives" );
        }

        // ---------------------------------------------------------------- H^1 metric and its factor
        // Weights of the H^1 metric (src/SimplicialMesh/H1Metric.hpp:4-5, setters :13-27).
    protected:
        double h1_stiffness_weight = 1.0;
        double h1_mass_weight      = 1.0;
        mutable std::unique_ptr<CSRMatrix> h1;
        mutable std::unique_ptr<SupernodalFactor> h1_factor;
        mutable bool h1_values_stale = false;   // h1_factor's structure is current, its values belong to older coordinates
        mutable size_t h1_pattern_nnz = 0;      // nonzeros of the H^1 pattern h1_factor was analysed for
        mutable std::unique_ptr<std::vector<Int>> nd_ordering;   // persistent (depends on the tree only)

        mutable Int_ h1_solver_id = 0;

    public:
        double H1StiffnessWeight() const override { return h1_stiffness_weight; }
        double H1MassWeight()      const override { return h1_mass_weight; }
        void SetH1StiffnessWeight( const double w ) override { h1_stiffness_weight = w; h1.reset(); if( h1_factor ) { h1_values_stale = true; } device_state.erase("h1"); }
        void SetH1MassWeight( const double w ) override      { h1_mass_weight = w;      h1.reset(); if( h1_factor ) { h1_values_stale = true; } device_state.erase("h1"); }

        // Ordering of the H^1 Cholesky factorisation (SimplicialMesh::SetH1SolverID, src/SimplicialMesh/H1Metric.hpp:174-257):
        // 0 = nested dissection on the cluster tree, the reference's default and the only ordering this
        // library builds; other IDs are refused with an error message and leave the ID unchanged.
        Int_ H1SolverID() const override { return h1_solver_id; }
        void SetH1SolverID( const Int_ ID ) const override
        {
            if( ID != 0 ) { eprint( ClassName() + "::SetH1SolverID: solver ID " + std::to_string( ID ) + " is not available; only 0 (nested dissection) is. Doing nothing." ); return; }
            h1_solver_id = ID;
        }

        const CSRMatrix & H1Metric() const
        {
            RequireHostCoordinates();
            if( !h1 )
            {
                h1 = std::make_unique<CSRMatrix>( AssembleH1Metric<DOM_DIM>( V_coords.data(), VertexCount(), simplices.data(), SimplexCount(),
                                                                             h1_stiffness_weight, h1_mass_weight ) );
            }
            return *h1;
        }

        const std::vector<Int> & NestedDissection() const
        {
            if( !nd_ordering ) { nd_ordering = std::make_unique<std::vector<Int>>( NestedDissectionOrdering<DOM_DIM>( GetClusterTree(), VertexCount() ) ); }
            return *nd_ordering;
        }

        const SupernodalFactor & H1Solver() const
        {
            if( h1_factor && h1_values_stale )
            {
                const CSRMatrix & H = H1Metric();
                if( H.ci.size() == h1_pattern_nnz && (Int)H.n == h1_factor->n )
                {
                    // same pattern: numeric refactorisation on the kept structure (bitwise a fresh factor)
                    SparseCholeskyNumeric( *h1_factor, H.rp, H.ci, H.v, threads );
                    h1_values_stale = false;
                }
                else { h1_factor.reset(); }
            }
            if( !h1_factor )
            {
                const CSRMatrix & H = H1Metric();
                h1_factor = std::make_unique<SupernodalFactor>( SparseCholeskyFactor( H.n, H.rp, H.ci, H.v, NestedDissection(), threads ) );
                h1_pattern_nnz = H.ci.size();
                h1_values_stale = false;
                dev_h1_structure = false;
            }
            return *h1_factor;
        }

        // ---------------------------------------------------------------- device
        // Uploads topology once and the per-state data after every SemiStaticUpdate.
        // The device context, created on first use (nothing uploaded yet).
        rp_ctx * Context() const
        {
            if( !dev ) { dev = std::make_unique<DeviceContext>( device ); dev_topology = false; dev_data = false; dev_h1_structure = false; }
            return dev->get();
        }

        rp_ctx * Device() const
        {
            rp_ctx * c = Context();
            if( !dev_topology )
            {
                const Tree_T & S = GetClusterTree();
                const BCT_T & B = GetBlockClusterTree();
                auto chk = [&]( int st, const char * w ) { if( st != RP_OK ) { throw std::runtime_error( std::string(w) + ": " + rp_last_error() ); } };
                chk( rp_upload_mesh( c, DOM_DIM, VertexCount(), SimplexCount(), simplices.data(), VertexCharges() ), "rp_upload_mesh" );
                chk( rp_upload_tree( c, S.ClusterCount(), S.C_left.data(), S.C_right.data(), S.C_begin.data(), S.C_end.data(),
                                     S.P_ordering.data(), S.pre_cols.data(), S.pre_vals.data() ), "rp_upload_tree" );
                const NearTiles & Tl = B.tiles;
                chk( rp_upload_near_tiles( c, Tl.leaf_count, Tl.leaf_begin.data(), Tl.leaf_end.data(), Tl.col_ptr.data(),
                                           Tl.col_prim.data(), Tl.col_mask.data() ), "rp_upload_near_tiles" );
                chk( rp_upload_far_pattern( c, B.far.rows, B.far.rp.data(), B.far.ci.data() ), "rp_upload_far_pattern" );
                const SymmetricCSR vf = B.VeryNearCSR();
                chk( rp_upload_verynear_pattern( c, vf.rows, vf.rp.data(), vf.ci.empty() ? nullptr : vf.ci.data(),
                                                 B.NearFieldSeparationParameter(), B.Settings().max_refinement ), "rp_upload_verynear_pattern" );
                dev_topology = true;
                dev_data = false;
            }
            if( !dev_data )
            {
                const Tree_T & S = GetClusterTree();
                auto st = rp_upload_data( c, V_coords.data(), S.P_near.data(), S.C_far.data() );
                if( st != RP_OK ) { throw std::runtime_error( std::string("rp_upload_data: ") + rp_last_error() ); }
                dev_data = true;
                device_state.clear();
            }
            return c;
        }

        bool DeviceReady() const { return dev != nullptr; }

        // Device context holding the H^1 factor of the current state (factorised on the host once
        // per state, like the reference's cached H1Solver).
        rp_ctx * DeviceH1() const
        {
            rp_ctx * c = Device();
            if( !device_state.count( "h1" ) )
            {
                const SupernodalFactor & F = H1Solver();
                if( dev_h1_structure )
                {
                    // the device already holds this factor's structure: new values only
                    auto st = rp_update_h1_values( c, F.vals.data() );
                    if( st != RP_OK ) { throw std::runtime_error( std::string("rp_update_h1_values: ") + rp_last_error() ); }
                }
                else
                {
                    std::vector<int32_t> perm( F.perm.begin(), F.perm.end() );
                    auto st = rp_upload_h1_factor( c, F.n, F.ns, perm.data(), F.sn_begin.data(), F.row_ptr.data(), F.rows.data(),
                                                   F.val_ptr.data(), F.vals.data(), F.sn_parent.data() );
                    if( st != RP_OK ) { throw std::runtime_error( std::string("rp_upload_h1_factor: ") + rp_last_error() ); }
                    dev_h1_structure = true;
                }
                device_state["h1"] = 1;
            }
            return c;
        }

        // Converts near metric values (width values per slot: 2 for the tangent-point metric, 1 for
        // the pseudo-Laplacian) from the device tile layout to the reference's CSR order (full
        // symmetric pattern of NearCSR()).
        void NearTileValuesToCSR( const double * tile_vals, const SymmetricCSR & A, double * csr_vals, const int width = 2 ) const
        {
            const NearTiles & Tl = GetBlockClusterTree().tiles;
            for( Int l = 0; l < Tl.leaf_count; ++l )
            {
                const Int ib = Tl.leaf_begin[(size_t)l];
                const Int nI = Tl.leaf_end[(size_t)l] - ib;
                for( Int i = ib; i < ib + nI; ++i )
                {
                    LInt pos = A.rp[(size_t)i];
                    LInt m = 0;
                    for( LInt k = Tl.col_ptr[(size_t)l]; k < Tl.col_ptr[(size_t)l + 1]; ++k, ++m )
                    {
                        if( (Tl.col_mask[(size_t)k] >> (i - ib)) & 1u )
                        {
                            const LInt slot = Tl.val_ptr[(size_t)l] + m * nI + (i - ib);
                            for( int w = 0; w < width; ++w ) { csr_vals[width * pos + w] = tile_vals[width * slot + w]; }
                            ++pos;
                        }
                    }
                }
            }
        }
    };

} // namespace Repulsor
