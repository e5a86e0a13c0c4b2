// extern "C" wrappers around the header-only host API (include/repulsor/), used by the Python
// package repulsor_amd and by bench.py. Handles are opaque pointers; status codes follow
// include/repulsor_hip.h.

#include "../../include/repulsor/Repulsor.hpp"

#include <cstdarg>

using namespace Repulsor;

namespace
{
    thread_local std::string g_err;

    struct MeshHandle
    {
        int n = 2;
        std::unique_ptr<SimplicialMesh<1>> m1;
        std::unique_ptr<SimplicialMesh<2>> m2;
    };

    struct OpHandle
    {
        double q = 6, p = 12;
        int precision = 0;
        int pc_precision = -1;
        double pc_weight = 0.0;
        std::vector<float> lp_params;
        int lp_nfeat = 0;
        bool lp_bf16 = true;
        int lp_guard_iter = 12;        // residual guard of learned solves (0: fall back only at max_iter)
        double lp_base_rate = 0.5;     // its reference rate, updated by fallback solves
    };

    template<typename F>
    int call( MeshHandle * h, F && f )
    {
        if( !h ) { g_err = "NULL mesh handle"; return RP_EINVAL; }
        try
        {
            if( h->n == 1 ) { return f( *h->m1 ); }
            return f( *h->m2 );
        }
        catch( const std::exception & e ) { g_err = e.what(); return RP_EINVAL; }
    }

    template<typename Mesh> TangentPointEnergy0<Mesh> make_energy( const OpHandle * o )
    {
        TangentPointEnergy0<Mesh> E( o->q, o->p );
        E.SetPrecision( (Precision)o->precision );
        return E;
    }

    template<typename Mesh> TangentPointMetric0<Mesh> make_metric( const OpHandle * o )
    {
        TangentPointMetric0<Mesh> A( o->q, o->p );
        A.SetPrecision( (Precision)o->precision );
        if( o->pc_precision >= 0 ) { A.SetPreconditionerPrecision( (Precision)o->pc_precision ); }
        A.SetPreconditionerDiagonalWeight( o->pc_weight );
        if( !o->lp_params.empty() ) { A.SetLearnedPreconditioner( o->lp_params, o->lp_nfeat, o->lp_bf16 ); }
        A.SetLearnedPreconditionerGuard( o->lp_guard_iter, o->lp_base_rate );
        return A;
    }

    int gen_out( const MeshData & M, double * coords, int32_t * simplices, int64_t * counts )
    {
        counts[0] = M.VertexCount();
        counts[1] = M.SimplexCount();
        if( coords )    { std::copy( M.coords.begin(), M.coords.end(), coords ); }
        if( simplices ) { std::copy( M.simplices.begin(), M.simplices.end(), simplices ); }
        return RP_OK;
    }
}

extern "C"
{

const char * rpm_last_error( void ) { return g_err.empty() ? rp_last_error() : g_err.c_str(); }

// ---------------------------------------------------------------- generators (two-call: sizes, then data)
int rpm_gen_torus( int Nu, int Nv, double R, double r, double sigma, uint64_t seed, double * coords, int32_t * simplices, int64_t * counts )
{
    return gen_out( MakeTorus( Nu, Nv, R, r, sigma, seed ), coords, simplices, counts );
}

int rpm_gen_icosphere( int level, int lmax, double amplitude, uint64_t seed, int threads, double * coords, int32_t * simplices, int64_t * counts )
{
    return gen_out( MakeIcosphere( level, lmax, amplitude, seed, threads ), coords, simplices, counts );
}

int rpm_gen_trefoil( int64_t N, int modes, double amplitude, uint64_t seed, double * coords, int32_t * simplices, int64_t * counts )
{
    return gen_out( MakeTrefoil( N, modes, amplitude, seed ), coords, simplices, counts );
}

// ---------------------------------------------------------------- mesh
void * rpm_mesh_new( int dom_dim, const double * coords, int64_t V, const int32_t * simplices, int64_t T, const double * charges, int threads )
{
    g_err.clear();
    try
    {
        auto * h = new MeshHandle();
        h->n = dom_dim;
        if( dom_dim == 1 ) { h->m1 = std::make_unique<SimplicialMesh<1>>( coords, V, simplices, T, threads, charges ); }
        else if( dom_dim == 2 ) { h->m2 = std::make_unique<SimplicialMesh<2>>( coords, V, simplices, T, threads, charges ); }
        else { delete h; g_err = "dom_dim must be 1 or 2"; return nullptr; }
        return h;
    }
    catch( const std::exception & e ) { g_err = e.what(); return nullptr; }
}

void rpm_mesh_free( void * h ) { delete (MeshHandle *)h; }

int rpm_mesh_settings( void * h, int split_threshold, double far_theta, double near_theta, int max_refinement, int device, int device_build )
{
    return call( (MeshHandle *)h, [&]( auto & M )
    {
        M.cluster_tree_settings.split_threshold = split_threshold;
        M.cluster_tree_settings.device_build = ( device_build != 0 );
        M.block_cluster_tree_settings.far_field_separation_parameter = far_theta;
        M.adaptivity_settings.theta = near_theta;
        M.adaptivity_settings.max_refinement = max_refinement;
        M.SetDevice( device );
        return RP_OK;
    });
}

// counts: V, T, C, L, depth, near nnz, far nnz, very near nnz, inter nnz, leaf blocks, tile columns, tile slots
int rpm_mesh_build( void * h, int64_t * counts )
{
    return call( (MeshHandle *)h, [&]( auto & M )
    {
        const auto & S = M.GetClusterTree();
        const auto & B = M.GetBlockClusterTree();
        if( counts )
        {
            counts[0] = M.VertexCount(); counts[1] = M.SimplexCount(); counts[2] = S.ClusterCount();
            counts[3] = S.LeafClusterCount(); counts[4] = S.TreeDepth();
            counts[5] = B.NearFieldInteractionCount(); counts[6] = B.FarFieldInteractionCount();
            counts[7] = B.VeryNearFieldInteractionCount(); counts[8] = B.PrimitiveIntersectionCount();
            counts[9] = (int64_t)B.leaf_blocks.size(); counts[10] = B.tiles.col_ptr.back(); counts[11] = B.tiles.val_ptr.back();
        }
        return RP_OK;
    });
}

int rpm_mesh_tree( void * h, int32_t * C_left, int32_t * C_right, int32_t * C_begin, int32_t * C_end,
                   int32_t * P_ordering, int32_t * P_inverse, int32_t * leaf_clusters )
{
    return call( (MeshHandle *)h, [&]( auto & M )
    {
        const auto & S = M.GetClusterTree();
        auto cp = []( const std::vector<Int> & v, int32_t * o ) { if( o ) { std::copy( v.begin(), v.end(), o ); } };
        cp( S.C_left, C_left ); cp( S.C_right, C_right ); cp( S.C_begin, C_begin ); cp( S.C_end, C_end );
        cp( S.P_ordering, P_ordering ); cp( S.P_inverse_ordering, P_inverse ); cp( S.leaf_clusters, leaf_clusters );
        return RP_OK;
    });
}

// P_near T x ND, P_far T x 10 (tree order), C_far C x 10, C_serialized C x 7, P_serialized T x (4 + 3(n+1)), pre_cols, pre_vals
int rpm_mesh_tree_data( void * h, double * P_near, double * P_far, double * C_far, double * C_ser, double * P_ser,
                        int32_t * pre_cols, double * pre_vals )
{
    return call( (MeshHandle *)h, [&]( auto & M )
    {
        const auto & S = M.GetClusterTree();
        auto cp = []( const auto & v, auto * o ) { if( o ) { std::copy( v.begin(), v.end(), o ); } };
        cp( S.P_near, P_near ); cp( S.P_far, P_far ); cp( S.C_far, C_far ); cp( S.C_serialized, C_ser ); cp( S.P_serialized, P_ser );
        cp( S.pre_cols, pre_cols ); cp( S.pre_vals, pre_vals );
        return RP_OK;
    });
}

// kind 0: near (full symmetric, reference layout), 1: far, 2: very near. Pass NULL arrays to query nnz via *nnz.
// For kind 0, row_end >= 0 keeps only the rows [0, row_end).
int rpm_mesh_pattern( void * h, int kind, int64_t * rp, int32_t * ci, int64_t * diag, int64_t * nnz, int32_t row_end )
{
    return call( (MeshHandle *)h, [&]( auto & M )
    {
        const auto & B = M.GetBlockClusterTree();
        SymmetricCSR A;
        if( kind == 0 ) { A = B.NearCSR( row_end ); }
        else if( kind == 1 ) { A = B.far; }
        else { A = B.VeryNearCSR(); }
        if( nnz ) { *nnz = A.NonzeroCount(); }
        if( rp )   { std::copy( A.rp.begin(), A.rp.end(), rp ); }
        if( ci )   { std::copy( A.ci.begin(), A.ci.end(), ci ); }
        if( diag ) { std::copy( A.diag.begin(), A.diag.end(), diag ); }
        return RP_OK;
    });
}

// Upper leaf blocks (I <= J) as 2 x nblocks int32; and the excluded (very near + inter) primitive pairs.
int rpm_mesh_leaf_blocks( void * h, int32_t * blocks, int32_t * vf_pairs, int32_t * inter_pairs )
{
    return call( (MeshHandle *)h, [&]( auto & M )
    {
        const auto & B = M.GetBlockClusterTree();
        if( blocks ) { for( size_t k = 0; k < B.leaf_blocks.size(); ++k ) { blocks[2*k] = B.leaf_blocks[k].first; blocks[2*k+1] = B.leaf_blocks[k].second; } }
        if( vf_pairs ) { for( size_t k = 0; k < B.verynear_pairs.size(); ++k ) { vf_pairs[2*k] = B.verynear_pairs[k].first; vf_pairs[2*k+1] = B.verynear_pairs[k].second; } }
        if( inter_pairs ) { for( size_t k = 0; k < B.inter_pairs.size(); ++k ) { inter_pairs[2*k] = B.inter_pairs[k].first; inter_pairs[2*k+1] = B.inter_pairs[k].second; } }
        return RP_OK;
    });
}

// Near tiles: col_ptr (leaves + 1), col_prim and col_mask (tile columns), val_ptr (leaves + 1).
int rpm_mesh_tiles( void * h, int64_t * col_ptr, int32_t * col_prim, uint32_t * col_mask, int64_t * val_ptr )
{
    return call( (MeshHandle *)h, [&]( auto & M )
    {
        const auto & Tl = M.GetBlockClusterTree().tiles;
        auto cp = []( const auto & v, auto * o ) { if( o ) { std::copy( v.begin(), v.end(), o ); } };
        cp( Tl.col_ptr, col_ptr ); cp( Tl.col_prim, col_prim ); cp( Tl.col_mask, col_mask ); cp( Tl.val_ptr, val_ptr );
        return RP_OK;
    });
}

// Device time of the last device build of the tree and block cluster tree (ms), or -1.
int rpm_mesh_build_device_ms( void * h, double * ms )
{
    return call( (MeshHandle *)h, [&]( auto & M )
    {
        *ms = -1.0;
        if( M.DeviceReady() ) { rp_last_timing( M.Context(), 14, ms ); }
        return RP_OK;
    });
}

int rpm_mesh_semistatic_update( void * h, const double * coords, int device_ptr )
{
    return call( (MeshHandle *)h, [&]( auto & M ) { M.SemiStaticUpdate( coords, false, device_ptr != 0 ); return RP_OK; } );
}

int rpm_mesh_max_safe_step( void * h, const double * U, double max_time, double TOL, int device_ptr, double * t, int64_t * stats )
{
    return call( (MeshHandle *)h, [&]( auto & M )
    {
        *t = M.MaximumSafeStepSize( U, max_time, TOL, false, device_ptr != 0 );
        if( stats ) { for( int k = 0; k < 3; ++k ) { stats[k] = M.CollisionStats()[k]; } }
        return RP_OK;
    });
}

// Device-computed rows of the last device-resident update (tree order), for comparison with the host's.
int rpm_mesh_device_state_rows( void * h, double * P_near, double * C_far )
{
    return call( (MeshHandle *)h, [&]( auto & M ) { return rp_download_state_rows( M.Device(), P_near, C_far, nullptr ); } );
}

int rpm_mesh_coords( void * h, double * coords )
{
    return call( (MeshHandle *)h, [&]( auto & M ) { const auto & c = M.VertexCoordinates(); std::copy( c.begin(), c.end(), coords ); return RP_OK; } );
}

// Ensures the device context holds the topology and current data (uploads if needed).
int rpm_mesh_device( void * h )
{
    return call( (MeshHandle *)h, [&]( auto & M ) { (void)M.Device(); return RP_OK; } );
}

int rpm_mesh_timing( void * h, int which, double * ms )
{
    return call( (MeshHandle *)h, [&]( auto & M ) { return rp_last_timing( M.Device(), which, ms ); } );
}

void * rpm_mesh_context( void * h )
{
    rp_ctx * c = nullptr;
    call( (MeshHandle *)h, [&]( auto & M ) { c = M.Device(); return RP_OK; } );
    return c;
}

// ---------------------------------------------------------------- operators (energy / metric parameters)
void * rpm_op_new( double q, double p, int precision )
{
    auto * o = new OpHandle();
    o->q = q; o->p = p; o->precision = precision;
    return o;
}

void rpm_op_free( void * o ) { delete (OpHandle *)o; }

// Precision of the pseudo-Laplacian inside the metric's preconditioner (-1: the operator's precision).
void rpm_op_set_pc_precision( void * o, int precision ) { ((OpHandle *)o)->pc_precision = precision; }

// Diagonal weight of the metric's preconditioner P = P0 + w diag(A)^-1 (0: reference P0, < 0: estimated).
void rpm_op_set_pc_weight( void * o, double w ) { ((OpHandle *)o)->pc_weight = w; }

// Residual guard of learned solves: check iteration (0: off) and reference rate per iteration.
void rpm_op_set_lp_guard( void * o, int check_iteration, double reference_rate )
{
    ((OpHandle *)o)->lp_guard_iter = check_iteration;
    ((OpHandle *)o)->lp_base_rate = reference_rate;
}

// Learned preconditioner model of the operator (nfeat = 0 removes it).
void rpm_op_set_learned( void * o, int nfeat, const float * params, int bf16 )
{
    auto * h = (OpHandle *)o;
    if( nfeat <= 0 || !params ) { h->lp_params.clear(); h->lp_nfeat = 0; return; }
    h->lp_params.assign( params, params + LearnedPreconditionerParameterCount( nfeat ) );
    h->lp_nfeat = nfeat;
    h->lp_bf16 = bf16 != 0;
}

int rpm_learned_feature_count( void ) { return LEARNED_PC_FEATURES; }

// Node features of the mesh's cluster tree (C x nfeat floats).
int rpm_learned_features( void * h, float * out )
{
    return call( (MeshHandle *)h, [&]( auto & M )
    {
        constexpr int n = std::decay_t<decltype(M)>::DOM_DIM;
        const auto f = LearnedPreconditionerFeatures<n>( M.GetClusterTree() );
        std::copy( f.begin(), f.end(), out );
        return RP_OK;
    });
}

int rpm_pc_learn_step( void * op, void * h, float lr, int step, float * params, float * m, float * v, double * loss, double * grad )
{
    return call( (MeshHandle *)h, [&]( auto & M )
    {
        using Mesh = std::decay_t<decltype(M)>;
        auto A = make_metric<Mesh>( (OpHandle *)op );
        return A.PreconditionerLearnStep( M, lr, step, params, m, v, loss, grad ) ? RP_OK : RP_EINVAL;
    });
}

int rpm_pc_scaling( void * op, void * h, double * sqrt_d )
{
    return call( (MeshHandle *)h, [&]( auto & M )
    {
        using Mesh = std::decay_t<decltype(M)>;
        return make_metric<Mesh>( (OpHandle *)op ).LearnedScaling( M, sqrt_d ) ? RP_OK : RP_EINVAL;
    });
}

// Weight in use on the mesh's device (after MultiplyPreconditioner / Solve).
int rpm_pc_weight( void * op, void * h, double * w )
{
    return call( (MeshHandle *)h, [&]( auto & M )
    {
        using Mesh = std::decay_t<decltype(M)>;
        *w = make_metric<Mesh>( (OpHandle *)op ).PreconditionerWeightInUse( M );
        return RP_OK;
    });
}

// Exact diagonal of the assembled metric (V values).
int rpm_metric_diagonal( void * op, void * h, double * diag )
{
    return call( (MeshHandle *)h, [&]( auto & M )
    {
        using Mesh = std::decay_t<decltype(M)>;
        return make_metric<Mesh>( (OpHandle *)op ).MetricDiagonal( M, diag ) ? RP_OK : RP_EINVAL;
    });
}

int rpm_energy_value( void * op, void * h, double * out )
{
    return call( (MeshHandle *)h, [&]( auto & M )
    {
        using Mesh = std::decay_t<decltype(M)>;
        auto E = make_energy<Mesh>( (OpHandle *)op );
        g_err.clear();
        *out = E.Value( M );
        return M.scalar_cache.count( E.ClassName() + "::Value" ) ? RP_OK : RP_EINVAL;
    });
}

int rpm_energy_differential( void * op, void * h, double alpha, double beta, double * Y, int64_t ldY, double * en, int device_ptr )
{
    return call( (MeshHandle *)h, [&]( auto & M )
    {
        rp_ctx * c = M.Device();
        int st = rp_tp0_compute( c, RP_ENERGY | RP_DIFF, ((OpHandle *)op)->q, ((OpHandle *)op)->p, ((OpHandle *)op)->precision, en );
        if( st != RP_OK ) { return st; }
        return rp_assemble_vertex_gradient( c, alpha, beta, Y, ldY, device_ptr ? RP_DEVICE_PTR : 0 );
    });
}

int rpm_energy_moment_derivatives( void * op, void * h, double * P_D_near, double * P_D_far )
{
    return call( (MeshHandle *)h, [&]( auto & M ) { (void)op; return rp_download_derivatives( M.Device(), P_D_near, P_D_far ); } );
}

int rpm_metric_values( void * op, void * h, double * nf_csr, double * nf_diag, double * ff_vals, double * ff_diag )
{
    return call( (MeshHandle *)h, [&]( auto & M )
    {
        using Mesh = std::decay_t<decltype(M)>;
        auto A = make_metric<Mesh>( (OpHandle *)op );
        return A.DownloadMetricValues( M, nf_csr, nf_diag, ff_vals, ff_diag ) ? RP_OK : RP_EINVAL;
    });
}

int rpm_vf_metric_values( void * op, void * h, double * vf_vals, double * vf_diag )
{
    return call( (MeshHandle *)h, [&]( auto & M )
    {
        using Mesh = std::decay_t<decltype(M)>;
        auto A = make_metric<Mesh>( (OpHandle *)op );
        return A.DownloadVeryNearMetricValues( M, vf_vals, vf_diag ) ? RP_OK : RP_EINVAL;
    });
}

int rpm_pl_vf_values( void * op, void * h, double * vf_vals, double * vf_diag )
{
    return call( (MeshHandle *)h, [&]( auto & M )
    {
        using Mesh = std::decay_t<decltype(M)>;
        auto A = make_metric<Mesh>( (OpHandle *)op );
        return A.DownloadVeryNearPseudoLaplacianValues( M, vf_vals, vf_diag ) ? RP_OK : RP_EINVAL;
    });
}

// Evaluations and deepest level of the very-near field of the mesh's last tangent-point compute.
int rpm_vf_stats( void * h, int64_t * evaluations, int * max_level )
{
    return call( (MeshHandle *)h, [&]( auto & M )
    {
        if( !M.DeviceReady() ) { *evaluations = 0; *max_level = 0; return RP_OK; }
        return rp_vf_stats( M.Device(), evaluations, max_level );
    });
}

int rpm_metric_multiply( void * op, void * h, double alpha, const double * X, int64_t ldX, double beta, double * Y, int64_t ldY,
                         int nrhs, int vf, int nf, int ff, int device_ptr )
{
    return call( (MeshHandle *)h, [&]( auto & M )
    {
        using Mesh = std::decay_t<decltype(M)>;
        auto A = make_metric<Mesh>( (OpHandle *)op );
        if( nrhs != 3 )
        {
            A.MultiplyMetric( M, alpha, X, ldX, beta, Y, ldY, nrhs, vf, nf, ff, device_ptr );
            g_err = "nrhs must be 3";
            return RP_EINVAL;
        }
        if( nf == RP_NF_FP32_COPIES ) { return A.MultiplyMetricInner( M, alpha, X, ldX, beta, Y, ldY, device_ptr ) ? RP_OK : RP_EINVAL; }
        return A.MultiplyMetric( M, alpha, X, ldX, beta, Y, ldY, nrhs, vf, nf, ff, device_ptr ) ? RP_OK : RP_EINVAL;
    });
}

int rpm_metric_assemble( void * op, void * h )
{
    return call( (MeshHandle *)h, [&]( auto & M )
    {
        using Mesh = std::decay_t<decltype(M)>;
        auto A = make_metric<Mesh>( (OpHandle *)op );
        return A.AssembleMetric( M ) ? RP_OK : RP_EINVAL;
    });
}

int rpm_metric_precondition( void * op, void * h, double alpha, const double * X, int64_t ldX, double beta, double * Y, int64_t ldY,
                             int nrhs, int device_ptr )
{
    return call( (MeshHandle *)h, [&]( auto & M )
    {
        using Mesh = std::decay_t<decltype(M)>;
        auto A = make_metric<Mesh>( (OpHandle *)op );
        const bool ok = A.MultiplyPreconditioner( M, alpha, X, ldX, beta, Y, ldY, nrhs, device_ptr );
        if( !ok && nrhs != 3 ) { g_err = "nrhs must be 3"; }
        return ok ? RP_OK : RP_EINVAL;
    });
}

// X = alpha A^+ B + beta X; iters (2: iteration count, 1 if the learned preconditioner fell back) and
// relres (3) receive the solver's record. Returns RP_ESTAGNATION
// when max_iter was reached before the tolerance.
int rpm_metric_solve( void * op, void * h, double alpha, const double * B, int64_t ldB, double beta, double * X, int64_t ldX,
                      int nrhs, int max_iter, double tol, int solver_flags, int device_ptr, int * iters, double * relres )
{
    return call( (MeshHandle *)h, [&]( auto & M )
    {
        using Mesh = std::decay_t<decltype(M)>;
        auto A = make_metric<Mesh>( (OpHandle *)op );
        (void)A.Solve( M, alpha, B, ldB, beta, X, ldX, nrhs, max_iter, tol, device_ptr, solver_flags );
        ((OpHandle *)op)->lp_base_rate = A.LearnedPreconditionerReferenceRate();
        const int st = A.Solver_Status();
        if( st != RP_OK && st != RP_ESTAGNATION ) { if( nrhs != 3 ) { g_err = "nrhs must be 3"; } return st; }
        const auto rr = A.Solver_RelativeResiduals();
        if( iters ) { iters[0] = A.Solver_IterationCount(); iters[1] = A.Solver_FellBack() ? 1 : 0; }
        bool conv = true;
        for( size_t k = 0; k < rr.size(); ++k ) { if( relres ) { relres[k] = rr[k]; } if( !( rr[k] <= tol ) ) { conv = false; } }
        return conv ? RP_OK : RP_ESTAGNATION;
    });
}

// Pseudo-Laplacian L_{2-s} of the metric's preconditioner: values in the reference's CSR order.
int rpm_pl_values( void * op, void * h, double * nf_csr, double * nf_diag, double * ff_vals, double * ff_diag )
{
    return call( (MeshHandle *)h, [&]( auto & M )
    {
        using Mesh = std::decay_t<decltype(M)>;
        auto A = make_metric<Mesh>( (OpHandle *)op );
        if( !A.PseudoLaplacianValues( M ) ) { return RP_EINVAL; }
        const auto & B = M.GetBlockClusterTree();
        std::vector<double> tile( (size_t)std::max<LInt>( 1, B.tiles.val_ptr.back() ) );
        const int st = rp_download_pl( M.Device(), nf_csr ? tile.data() : nullptr, nf_diag, ff_vals, ff_diag );
        if( st != RP_OK ) { g_err = rp_last_error(); return st; }
        if( nf_csr ) { M.NearTileValuesToCSR( tile.data(), B.NearCSR(), nf_csr, 1 ); }
        return RP_OK;
    });
}

int rpm_pl_multiply( void * op, void * h, double alpha, const double * X, int64_t ldX, double beta, double * Y, int64_t ldY, int device_ptr )
{
    return call( (MeshHandle *)h, [&]( auto & M )
    {
        using Mesh = std::decay_t<decltype(M)>;
        auto A = make_metric<Mesh>( (OpHandle *)op );
        if( !A.PseudoLaplacianValues( M ) ) { return RP_EINVAL; }
        const int st = rp_pl_multiply( M.Device(), alpha, X, ldX, beta, Y, ldY, 3, 1, 1, device_ptr ? RP_DEVICE_PTR : 0 );
        if( st != RP_OK ) { g_err = rp_last_error(); }
        return st;
    });
}

// X = alpha H^-1 B + beta X with the H^1 factor on the device.
int rpm_h1_solve_device( void * h, double alpha, const double * B, int64_t ldB, double beta, double * X, int64_t ldX, int device_ptr )
{
    return call( (MeshHandle *)h, [&]( auto & M )
    {
        rp_ctx * c = M.DeviceH1();
        const int st = rp_h1_solve( c, alpha, B, ldB, beta, X, ldX, 3, device_ptr ? RP_DEVICE_PTR : 0 );
        if( st != RP_OK ) { g_err = rp_last_error(); }
        return st;
    });
}

// H^1 metric factor on the host: x = H^{-1} b (V x 3); stats: nnz(H), nnz(L), supernodes, levels, flops.
int rpm_h1_solve_host( void * h, const double * b, double * x, double * stats )
{
    return call( (MeshHandle *)h, [&]( auto & M )
    {
        const auto & F = M.H1Solver();
        if( b && x ) { F.Solve( b, x ); }
        if( stats )
        {
            stats[0] = (double)M.H1Metric().ci.size(); stats[1] = (double)F.nnz_L; stats[2] = F.ns;
            stats[3] = F.levels; stats[4] = F.flops;
        }
        return RP_OK;
    });
}

// y = H x (V x 3), the assembled H^1 metric.
int rpm_h1_multiply_host( void * h, const double * x, double * y )
{
    return call( (MeshHandle *)h, [&]( auto & M ) { M.H1Metric().Multiply3( x, y ); return RP_OK; } );
}

// The nested-dissection ordering (new -> old).
int rpm_nd_ordering( void * h, int32_t * perm )
{
    return call( (MeshHandle *)h, [&]( auto & M ) { const auto & p = M.NestedDissection(); std::copy( p.begin(), p.end(), perm ); return RP_OK; } );
}

// H^1 metric in CSR (pass NULL arrays to query nnz).
int rpm_h1_csr( void * h, int64_t * rp, int32_t * ci, double * v, int64_t * nnz )
{
    return call( (MeshHandle *)h, [&]( auto & M )
    {
        const auto & H = M.H1Metric();
        if( nnz ) { *nnz = (int64_t)H.ci.size(); }
        if( rp ) { std::copy( H.rp.begin(), H.rp.end(), rp ); }
        if( ci ) { std::copy( H.ci.begin(), H.ci.end(), ci ); }
        if( v )  { std::copy( H.v.begin(), H.v.end(), v ); }
        return RP_OK;
    });
}

// Host timers of the header-only API (name \t seconds lines).
int rpm_timer_report( char * buf, int64_t cap )
{
    std::string s;
    auto & tt = TimerTable::Get();
    std::lock_guard<std::mutex> lk( tt.mtx );
    for( auto & r : tt.rows ) { s += r.first + "\t" + std::to_string( r.second ) + "\n"; }
    if( buf && cap > 0 ) { std::snprintf( buf, (size_t)cap, "%s", s.c_str() ); }
    return (int)s.size();
}

} // extern "C"
