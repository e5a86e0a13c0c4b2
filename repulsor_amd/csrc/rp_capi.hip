// C-ABI entry points of librepulsor_hip.so (declared in include/repulsor_hip.h).

#include "rp_capi_util.hpp"
#include "k_pair.hpp"
#include "rp_launch.hpp"

#include <cstring>
#include <mutex>

namespace rp
{
    // Scratch of rp_tp0_energy_noalloc: one array per device (each device holds its own copy until the
    // process ends). Host code never passes the array's name to a kernel or a copy: it uses the address
    // that hipGetSymbolAddress returns on each device (scratch_address), looked up once per device when
    // a context is created there. Every call on a device shares the array: overlapping calls on one
    // device race.
    constexpr int RP_SCRATCH_DEVICES = 64;
    static __device__ __align__(256) double scratch[RP_NOALLOC_SCRATCH_DOUBLES];
    static std::mutex scratch_mutex;
    static double * scratch_ptr[RP_SCRATCH_DEVICES] = {};

    // Looks up (once) the address of `scratch` on the current device `device`.
    static void scratch_setup( int device )
    {
        if( device < 0 || device >= RP_SCRATCH_DEVICES ) { return; }
        std::lock_guard<std::mutex> lock( scratch_mutex );
        if( scratch_ptr[device] ) { return; }
        void * p = nullptr;
        RP_HIP( hipGetSymbolAddress( &p, HIP_SYMBOL(scratch) ) );
        require( p != nullptr, "hipGetSymbolAddress returned no address for the scratch array", RP_EHIP );
        scratch_ptr[device] = static_cast<double *>( p );
    }

    static double * scratch_address( int device )
    {
        if( device < 0 || device >= RP_SCRATCH_DEVICES ) { return nullptr; }
        std::lock_guard<std::mutex> lock( scratch_mutex );
        return scratch_ptr[device];
    }

    static thread_local std::string g_last_error;
    void set_error( const std::string & msg ) { g_last_error = msg; }

    static int exponent_kind( double q, double p )
    {
        if( q == 6.0 && p == 12.0 ) { return KIND_Q6P12; }
        if( q == 2.0 && p == 4.0 )  { return KIND_Q2P4; }
        return KIND_GENERIC;
    }

    // Y = alpha A X + beta Y on device vectors (V x 3, leading dimensions ldx, ldy): pre-operator,
    // cluster sums, far-field SpMM and down-sweep, near-field SpMM with fused post epilogue, vertex gather.
    // inner32: with a metric assembled at RP_MIXED_REFINED, the near field multiplies with the FP32
    // copies and their diagonal (far and very-near fields stay FP64).
    // Reference: MetricDimRestricted::MultiplyMetric (src/Metrics/MetricDimRestricted.hpp:109-139).
    void metric_apply( rp_ctx * ctx, double alpha, const double * Xd, int64_t ldx, double beta, double * Yd, int64_t ldy,
                       bool nf_flag, bool ff_flag, bool time_nf, bool vf_flag, bool inner32 )
    {
        const int64_t T = ctx->T, V = ctx->V, C = ctx->C, L = ctx->L;
        const bool mixed = ctx->metric_precision == RP_MIXED;
        const bool use32 = inner32 && ctx->metric_precision == RP_MIXED_REFINED;
        ctx->P_in.require( (size_t)(12 * T) ); ctx->P_out.require( (size_t)(12 * T) );
        ctx->C_in.require( (size_t)(12 * C) ); ctx->C_out.require( (size_t)(12 * C) );
        ctx->leaf_out.require( (size_t)(12 * L) );
        const bool do_nf = nf_flag && ctx->nf_cols > 0;
        const bool do_ff = ff_flag && ctx->far_nnz > 0;
        launch_pre( ctx->size, T, ctx->pre_cols.get(), ctx->pre_vals.get(), Xd, ldx, ctx->P_in.get(), ctx->stream );
        if( do_ff )
        {
            launch_up_leaves( L, ctx->leaf_cluster.get(), ctx->C_begin.get(), ctx->C_end.get(), ctx->P_in.get(), ctx->C_in.get(), ctx->stream );
            launch_up_tree( ctx->h_level_ptr, ctx->level_nodes.get(), ctx->C_left.get(), ctx->C_right.get(), ctx->C_in.get(), ctx->stream, 12 );
            // Far field first, so that the near-field SpMM epilogue can add the percolated cluster
            // output of its leaf and apply the post-operator (S_out = per-vertex-slot contributions).
            launch_ff_spmm( mixed, C, ctx->far_rp.get(), ctx->far_ci.get(), ctx->ff_vals.get(), ctx->ff_diag.get(), ctx->C_in.get(), ctx->C_out.get(), ctx->stream );
            launch_leaf_out( L, C, ctx->leaf_anc.get(), ctx->leaf_anc_len.get(), ctx->C_out.get(), ctx->leaf_out.get(), ctx->stream );
        }
        double * S_out = ctx->P_out.get();   // T x (n+1) x 3 <= 12 T
        if( time_nf ) { RP_HIP( hipEventRecord( ctx->ev[6][0], ctx->stream ) ); }
        if( do_nf )
        {
            launch_nf_spmm_post( mixed || use32, ctx->size, T, L, ctx->leaf_begin.get(), ctx->leaf_end.get(), ctx->col_ptr.get(), ctx->col_prim.get(),
                                 ctx->col_mask.get(), ctx->val_ptr.get(), use32 ? ctx->nf_vals32.get() : ctx->nf_vals.get(),
                                 use32 ? ctx->nf_diag32.get() : ctx->nf_diag.get(), ctx->P_in.get(),
                                 ctx->leaf_out.get(), do_ff ? 1 : 0, ctx->pre_vals.get(), S_out, ctx->stream );
        }
        else
        {
            launch_post_prim( ctx->size, T, ctx->leaf_of_prim.get(), ctx->pre_vals.get(), ctx->leaf_out.get(), do_ff ? 1 : 0, S_out, ctx->stream );
        }
        if( vf_flag && ctx->have_vf_metric )
        {
            launch_vf_spmm_post( ctx->size, ctx->vf_nrows, ctx->vf_rows.get(), ctx->vf_rp.get(), ctx->vf_ci.get(), ctx->vf_vals.get(),
                                 ctx->vf_diag.get(), ctx->P_in.get(), ctx->pre_vals.get(), S_out, ctx->stream );
        }
        if( time_nf ) { RP_HIP( hipEventRecord( ctx->ev[6][1], ctx->stream ) ); }
        launch_post_vert( V, ctx->vp_ptr.get(), ctx->vp_entry.get(), S_out, alpha, beta, Yd, ldy, ctx->stream );
    }

    // Leaf-path sums of the far-field derivatives (needed by both read-out paths).
    static void collect_far( rp_ctx * ctx )
    {
        ctx->leaf_far.require( (size_t)(10 * ctx->L) );
        launch_leaf_far( ctx->L, ctx->C, ctx->leaf_anc.get(), ctx->leaf_anc_len.get(), ctx->ff_dx.get(), ctx->leaf_far.get(), ctx->stream );
    }
}

using namespace rp;

extern "C"
{

const char * rp_last_error(void) { return g_last_error.c_str(); }

const char * rp_version(void) { return "repulsor_hip 0.1 (gfx950)"; }

int rp_ctx_create( int device, rp_ctx ** out )
{
    return guarded( "rp_ctx_create", [&]()
    {
        require( out != nullptr, "out is NULL" );
        *out = nullptr;
        int count = 0;
        RP_HIP( hipGetDeviceCount( &count ) );
        require( device >= 0 && device < count, "no such HIP device", RP_EHIP );
        RP_HIP( hipSetDevice( device ) );
        scratch_setup( device );
        rp_ctx * c = new rp_ctx();
        c->device = device;
        try
        {
            RP_HIP( hipStreamCreateWithFlags( &c->stream, hipStreamNonBlocking ) );
            for( int k = 0; k < 16; ++k ) { RP_HIP( hipEventCreate( &c->ev[k][0] ) ); RP_HIP( hipEventCreate( &c->ev[k][1] ) ); }
            RP_HIP( hipEventCreateWithFlags( &c->ev_check, hipEventDisableTiming ) );
            RP_HIP( hipHostMalloc( (void **)&c->h_pinned, 16 * sizeof(double), hipHostMallocDefault ) );
        }
        catch( ... ) { delete c; throw; }
        *out = c;
    });
}

int rp_ctx_destroy( rp_ctx * ctx )
{
    if( !ctx ) { return RP_OK; }
    (void)hipSetDevice( ctx->device );
    if( ctx->stream ) { (void)hipStreamSynchronize( ctx->stream ); }
    for( int k = 0; k < 16; ++k ) { (void)hipEventDestroy( ctx->ev[k][0] ); (void)hipEventDestroy( ctx->ev[k][1] ); }
    if( ctx->ev_check ) { (void)hipEventDestroy( ctx->ev_check ); }
    if( ctx->h_pinned ) { (void)hipHostFree( ctx->h_pinned ); }
    if( ctx->stream ) { (void)hipStreamDestroy( ctx->stream ); }
    delete ctx;
    return RP_OK;
}

int rp_upload_mesh( rp_ctx * ctx, int dom_dim, int64_t vertex_count, int64_t simplex_count,
                    const int32_t * simplices, const double * vertex_charges )
{
    return guarded( "rp_upload_mesh", [&]()
    {
        require( ctx != nullptr, "ctx is NULL" );
        require( dom_dim == 1 || dom_dim == 2, "only curves (dom_dim 1) and surfaces (dom_dim 2) in R^3 are supported", RP_EUNSUPPORTED );
        require( vertex_count > 0 && simplex_count > 0 && simplices != nullptr, "empty mesh" );
        require( vertex_count < (int64_t)INT32_MAX && simplex_count * (dom_dim + 1) < (int64_t)INT32_MAX, "mesh too large for 32-bit indices" );
        RP_HIP( hipSetDevice( ctx->device ) );
        const int size = dom_dim + 1;
        for( int64_t r = 0; r < simplex_count * size; ++r )
        {
            require( simplices[r] >= 0 && simplices[r] < vertex_count, "simplex vertex index out of range" );
        }
        ctx->dom_dim = dom_dim;
        ctx->size = size;
        ctx->nd = 1 + 3 * size + 6;
        ctx->V = vertex_count;
        ctx->T = simplex_count;
        ctx->simplices.upload( simplices, (size_t)(simplex_count * size), ctx->stream );
        ctx->have_charges = ( vertex_charges != nullptr );
        if( vertex_charges ) { ctx->charges.upload( vertex_charges, (size_t)vertex_count, ctx->stream ); }
        // vertex -> corner CSR, corners increasing
        std::vector<int64_t> ptr( (size_t)vertex_count + 1, 0 );
        for( int64_t r = 0; r < simplex_count * size; ++r ) { ++ptr[(size_t)simplices[r] + 1]; }
        for( int64_t v = 0; v < vertex_count; ++v ) { ptr[(size_t)v + 1] += ptr[(size_t)v]; }
        std::vector<int32_t> corner( (size_t)(simplex_count * size) );
        std::vector<int64_t> pos( ptr.begin(), ptr.end() - 1 );
        for( int64_t r = 0; r < simplex_count * size; ++r ) { corner[(size_t)pos[(size_t)simplices[r]]++] = (int32_t)r; }
        ctx->vc_ptr.upload( ptr.data(), ptr.size(), ctx->stream );
        ctx->vc_corner.upload( corner.data(), corner.size(), ctx->stream );
        RP_HIP( hipStreamSynchronize( ctx->stream ) );
        ctx->have_tree = ctx->have_near = ctx->have_far = ctx->have_data = false;
        ctx->have_diff = ctx->have_metric = ctx->have_pl = ctx->have_h1 = ctx->have_h1_structure = false;
    });
}

int rp_upload_tree( rp_ctx * ctx, int32_t cluster_count,
                    const int32_t * C_left, const int32_t * C_right,
                    const int32_t * C_begin, const int32_t * C_end,
                    const int32_t * P_ordering,
                    const int32_t * pre_cols, const double * pre_vals )
{
    return guarded( "rp_upload_tree", [&]()
    {
        require( ctx != nullptr && ctx->T > 0, "upload the mesh first", RP_ENOTREADY );
        require( cluster_count > 0 && C_left && C_right && C_begin && C_end && P_ordering && pre_cols && pre_vals, "missing tree arrays" );
        RP_HIP( hipSetDevice( ctx->device ) );
        const int64_t C = cluster_count, T = ctx->T;
        const int size = ctx->size;
        std::vector<int32_t> parent( (size_t)C, -1 ), depth( (size_t)C, 0 );
        std::vector<int32_t> leaves;
        for( int64_t c = 0; c < C; ++c )
        {
            require( C_begin[c] >= 0 && C_end[c] <= T && C_begin[c] < C_end[c], "bad cluster range" );
            const int32_t L = C_left[c], R = C_right[c];
            if( L >= 0 )
            {
                require( L > c && R > c && L < C && R < C, "children must follow their parent in depth-first numbering" );
                parent[(size_t)L] = (int32_t)c; parent[(size_t)R] = (int32_t)c;
            }
            else { leaves.push_back( (int32_t)c ); }
        }
        for( int64_t c = 0; c < C; ++c ) { if( parent[(size_t)c] >= 0 ) { depth[(size_t)c] = depth[(size_t)parent[(size_t)c]] + 1; } }
        int maxd = 0;
        for( auto d : depth ) { maxd = std::max( maxd, d ); }
        require( maxd < 90, "cluster tree deeper than 90 levels" );
        const int64_t L = (int64_t)leaves.size();
        std::vector<int32_t> leaf_of_prim( (size_t)T, -1 );
        int max_leaf = 0;
        for( int64_t l = 0; l < L; ++l )
        {
            const int32_t c = leaves[(size_t)l];
            max_leaf = std::max( max_leaf, C_end[c] - C_begin[c] );
            for( int32_t i = C_begin[c]; i < C_end[c]; ++i ) { leaf_of_prim[(size_t)i] = (int32_t)l; }
        }
        for( auto x : leaf_of_prim ) { require( x >= 0, "leaves do not cover all primitives" ); }
        for( int64_t i = 0; i < T; ++i ) { require( P_ordering[i] >= 0 && P_ordering[i] < T, "bad P_ordering" ); }
        // root-to-leaf paths, root first, depth-major so that the leaves of a wavefront read one row
        std::vector<int32_t> anc( (size_t)( maxd + 1 ) * (size_t)L ), anc_len( (size_t)L );
        for( int64_t l = 0; l < L; ++l )
        {
            int32_t c = leaves[(size_t)l];
            const int n = depth[(size_t)c] + 1;
            anc_len[(size_t)l] = n;
            for( int d = n - 1; d >= 0; --d, c = parent[(size_t)c] ) { anc[(size_t)d * (size_t)L + (size_t)l] = c; }
        }
        // interior clusters grouped by depth, deepest first
        std::vector<int32_t> level_ptr( (size_t)maxd + 2, 0 );
        for( int64_t c = 0; c < C; ++c ) { if( C_left[c] >= 0 ) { ++level_ptr[(size_t)(maxd - depth[(size_t)c]) + 1]; } }
        for( int d = 0; d <= maxd; ++d ) { level_ptr[(size_t)d + 1] += level_ptr[(size_t)d]; }
        std::vector<int32_t> level_nodes( (size_t)std::max<int32_t>( 1, level_ptr.back() ) );
        {
            std::vector<int32_t> pos( level_ptr.begin(), level_ptr.end() - 1 );
            for( int64_t c = 0; c < C; ++c ) { if( C_left[c] >= 0 ) { level_nodes[(size_t)pos[(size_t)(maxd - depth[(size_t)c])]++] = (int32_t)c; } }
        }
        // vertex -> (tree primitive, slot) entries, increasing primitive
        const int64_t V = ctx->V;
        std::vector<int64_t> vptr( (size_t)V + 1, 0 );
        for( int64_t e = 0; e < T * size; ++e )
        {
            require( pre_cols[e] >= 0 && pre_cols[e] < V, "bad pre_cols" );
            ++vptr[(size_t)pre_cols[e] + 1];
        }
        for( int64_t v = 0; v < V; ++v ) { vptr[(size_t)v + 1] += vptr[(size_t)v]; }
        std::vector<int32_t> ventry( (size_t)(T * size) );
        {
            std::vector<int64_t> pos( vptr.begin(), vptr.end() - 1 );
            for( int64_t e = 0; e < T * size; ++e ) { ventry[(size_t)pos[(size_t)pre_cols[e]]++] = (int32_t)e; }
        }
        ctx->C = C; ctx->L = L; ctx->max_depth = maxd; ctx->max_leaf = max_leaf;
        ctx->C_left.upload( C_left, (size_t)C, ctx->stream );
        ctx->C_right.upload( C_right, (size_t)C, ctx->stream );
        ctx->C_begin.upload( C_begin, (size_t)C, ctx->stream );
        ctx->C_end.upload( C_end, (size_t)C, ctx->stream );
        ctx->C_parent.upload( parent.data(), (size_t)C, ctx->stream );
        ctx->leaf_anc.upload( anc.data(), anc.size(), ctx->stream );
        ctx->leaf_anc_len.upload( anc_len.data(), anc_len.size(), ctx->stream );
        ctx->P_ordering.upload( P_ordering, (size_t)T, ctx->stream );
        ctx->leaf_of_prim.upload( leaf_of_prim.data(), (size_t)T, ctx->stream );
        ctx->leaf_cluster.upload( leaves.data(), (size_t)L, ctx->stream );
        ctx->pre_cols.upload( pre_cols, (size_t)(T * size), ctx->stream );
        ++ctx->tree_version;
        ctx->pre_vals.upload( pre_vals, (size_t)(T * 4 * size), ctx->stream );
        ctx->vp_ptr.upload( vptr.data(), vptr.size(), ctx->stream );
        ctx->vp_entry.upload( ventry.data(), ventry.size(), ctx->stream );
        ctx->h_level_ptr = level_ptr;
        ctx->level_nodes.upload( level_nodes.data(), level_nodes.size(), ctx->stream );
        RP_HIP( hipStreamSynchronize( ctx->stream ) );
        ctx->have_tree = true;
        ctx->have_near = ctx->have_far = ctx->have_data = ctx->have_diff = ctx->have_metric = ctx->have_pl = false;
        ctx->vf_nnz = 0; ctx->vf_nrows = 0;
        ctx->have_vf_dh = ctx->have_vf_metric = ctx->have_pl_vf = false;
    });
}

int rp_upload_near_tiles( rp_ctx * ctx, int32_t leaf_count,
                          const int32_t * leaf_begin, const int32_t * leaf_end,
                          const int64_t * col_ptr, const int32_t * col_prim, const uint32_t * col_mask )
{
    return guarded( "rp_upload_near_tiles", [&]()
    {
        require( ctx != nullptr && ctx->have_tree, "upload the tree first", RP_ENOTREADY );
        require( leaf_count == ctx->L, "leaf_count differs from the tree's leaf count" );
        require( leaf_begin && leaf_end && col_ptr, "missing tile arrays" );
        RP_HIP( hipSetDevice( ctx->device ) );
        const int64_t L = leaf_count;
        std::vector<int64_t> val_ptr( (size_t)L + 1, 0 );
        require( col_ptr[0] == 0, "col_ptr[0] must be 0" );
        for( int64_t l = 0; l < L; ++l )
        {
            const int64_t n = leaf_end[l] - leaf_begin[l];
            require( n >= 1 && n <= 32, "leaves must hold 1..32 primitives (split_threshold <= 32)", RP_EUNSUPPORTED );
            require( leaf_begin[l] >= 0 && leaf_end[l] <= ctx->T, "bad leaf range" );
            require( col_ptr[l + 1] >= col_ptr[l], "col_ptr must be non-decreasing" );
            val_ptr[(size_t)l + 1] = val_ptr[(size_t)l] + n * ( col_ptr[l + 1] - col_ptr[l] );
        }
        const int64_t M = col_ptr[L];
        require( M == 0 || (col_prim && col_mask), "missing column arrays" );
        for( int64_t k = 0; k < M; ++k ) { require( col_prim[k] >= 0 && col_prim[k] < ctx->T, "column primitive out of range" ); }
        ctx->leaf_begin.upload( leaf_begin, (size_t)L, ctx->stream );
        ctx->leaf_end.upload( leaf_end, (size_t)L, ctx->stream );
        ctx->col_ptr.upload( col_ptr, (size_t)L + 1, ctx->stream );
        ctx->val_ptr.upload( val_ptr.data(), (size_t)L + 1, ctx->stream );
        if( M > 0 )
        {
            ctx->col_prim.upload( col_prim, (size_t)M, ctx->stream );
            ctx->col_mask.upload( col_mask, (size_t)M, ctx->stream );
        }
        ctx->nf_cols = M;
        ctx->nf_slots = val_ptr.back();
        RP_HIP( hipStreamSynchronize( ctx->stream ) );
        ctx->have_near = true;
        ctx->have_diff = ctx->have_metric = ctx->have_pl = false;
    });
}

int rp_upload_far_pattern( rp_ctx * ctx, int32_t rows, const int64_t * rp_, const int32_t * ci )
{
    return guarded( "rp_upload_far_pattern", [&]()
    {
        require( ctx != nullptr && ctx->have_tree, "upload the tree first", RP_ENOTREADY );
        require( rows == ctx->C && rp_ != nullptr, "far pattern must have one row per cluster" );
        RP_HIP( hipSetDevice( ctx->device ) );
        const int64_t nnz = rp_[rows];
        require( rp_[0] == 0 && nnz >= 0, "bad row pointers" );
        for( int64_t k = 0; k < nnz; ++k ) { require( ci[k] >= 0 && ci[k] < rows, "far column out of range" ); }
        ctx->far_rp.upload( rp_, (size_t)rows + 1, ctx->stream );
        if( nnz > 0 ) { ctx->far_ci.upload( ci, (size_t)nnz, ctx->stream ); }
        ctx->far_nnz = nnz;
        RP_HIP( hipStreamSynchronize( ctx->stream ) );
        ctx->have_far = true;
        ctx->have_diff = ctx->have_metric = ctx->have_pl = false;
    });
}

int rp_upload_verynear_pattern( rp_ctx * ctx, int32_t rows, const int64_t * rp_, const int32_t * ci, double theta, int max_refinement )
{
    return guarded( "rp_upload_verynear_pattern", [&]()
    {
        require( ctx != nullptr && ctx->have_tree, "upload the tree first", RP_ENOTREADY );
        require( rows == ctx->T && rp_ != nullptr, "very-near pattern must have one row per simplex" );
        require( std::isfinite( theta ) && theta > 0.0, "theta must be positive" );
        require( max_refinement >= 1, "max_refinement must be at least 1" );
        RP_HIP( hipSetDevice( ctx->device ) );
        const int64_t T = rows, nnz = rp_[rows];
        require( rp_[0] == 0 && nnz >= 0, "bad row pointers" );
        std::vector<int64_t> dg( (size_t)T ), tr( (size_t)std::max<int64_t>( 1, nnz ) );
        std::vector<int32_t> row( (size_t)std::max<int64_t>( 1, nnz ) ), rws;
        for( int64_t i = 0; i < T; ++i )
        {
            require( rp_[i + 1] >= rp_[i], "row pointers must be non-decreasing" );
            dg[(size_t)i] = rp_[i + 1];
            for( int64_t k = rp_[i]; k < rp_[i + 1]; ++k )
            {
                require( ci[k] >= 0 && ci[k] < T && ci[k] != i, "very-near column out of range or on the diagonal" );
                require( k == rp_[i] || ci[k] > ci[k - 1], "very-near columns must be strictly increasing per row" );
                if( ci[k] > i && dg[(size_t)i] == rp_[i + 1] ) { dg[(size_t)i] = k; }
                row[(size_t)k] = (int32_t)i;
            }
            if( rp_[i + 1] > rp_[i] ) { rws.push_back( (int32_t)i ); }
        }
        // metric values of a lower entry (i, j), j < i, come from the upper entry (j, i)
        for( int64_t i = 0; i < T; ++i )
        {
            for( int64_t k = rp_[i]; k < rp_[i + 1]; ++k )
            {
                const int32_t j = ci[k];
                if( j > i ) { tr[(size_t)k] = k; continue; }
                const int32_t * b = ci + dg[(size_t)j], * e = ci + rp_[j + 1];
                const int32_t * f = std::lower_bound( b, e, (int32_t)i );
                require( f != e && *f == (int32_t)i, "very-near pattern is not symmetric" );
                tr[(size_t)k] = f - ci;
            }
        }
        // deepest level: the reference clamps max_refinement to the largest level whose path code
        // fits a 64-bit integer (FMM_Kernel_VF.hpp:178-183, SimplexHierarchy_{1D,2D}.hpp MaxLevel)
        const double children = ( ctx->dom_dim == 2 ) ? 4.0 : 2.0;
        const int max_level = (int)std::floor( std::log( (double)std::numeric_limits<uint64_t>::max() ) / std::log( children ) );
        ctx->vf_max_ref = std::min( max_refinement, max_level );
        ctx->vf_theta = theta;
        ctx->vf_rp.upload( rp_, (size_t)T + 1, ctx->stream );
        ctx->vf_dg.upload( dg.data(), (size_t)T, ctx->stream );
        if( nnz > 0 )
        {
            ctx->vf_ci.upload( ci, (size_t)nnz, ctx->stream );
            ctx->vf_row.upload( row.data(), (size_t)nnz, ctx->stream );
            ctx->vf_tr.upload( tr.data(), (size_t)nnz, ctx->stream );
            ctx->vf_rows.upload( rws.data(), rws.size(), ctx->stream );
        }
        RP_HIP( hipStreamSynchronize( ctx->stream ) );
        ctx->vf_nnz = nnz;
        ctx->vf_nrows = (int64_t)rws.size();
        ctx->have_diff = ctx->have_metric = ctx->have_pl = false;
        ctx->have_vf_dh = ctx->have_vf_metric = ctx->have_pl_vf = false;
    });
}

int rp_build_topology( rp_ctx * ctx, int dom_dim, int64_t vertex_count, int64_t simplex_count, const double * V_coords,
                       const int32_t * simplices, int split_threshold, double far_theta, double near_theta, double intersection_theta,
                       int64_t * counts )
{
    return guarded( "rp_build_topology", [&]()
    {
        require( ctx != nullptr, "ctx is NULL" );
        require( dom_dim == 1 || dom_dim == 2, "only curves (dom_dim 1) and surfaces (dom_dim 2) in R^3 are supported", RP_EUNSUPPORTED );
        require( vertex_count > 0 && simplex_count > 0 && V_coords && simplices, "empty mesh" );
        require( vertex_count < (int64_t)INT32_MAX && 2 * simplex_count * ( dom_dim + 1 ) < (int64_t)INT32_MAX, "mesh too large for 32-bit indices" );
        require( split_threshold >= 1 && split_threshold <= 32, "split_threshold must lie in 1..32", RP_EUNSUPPORTED );
        require( std::isfinite( far_theta ) && far_theta >= 0.0 && std::isfinite( near_theta ) && near_theta >= 0.0
                 && std::isfinite( intersection_theta ) && intersection_theta >= 0.0, "separation parameters must be finite and non-negative" );
        const int size = dom_dim + 1;
        for( int64_t r = 0; r < simplex_count * size; ++r )
        {
            require( simplices[r] >= 0 && simplices[r] < vertex_count, "simplex vertex index out of range" );
        }
        for( int64_t r = 0; r < 3 * vertex_count; ++r ) { require( std::isfinite( V_coords[r] ), "coordinates contain NaN or Inf" ); }
        RP_HIP( hipSetDevice( ctx->device ) );
        ctx->built.reset();
        auto B = std::make_unique<BuiltTopology>();
        EventScope ev( ctx, 14 );
        build_topology( *B, dom_dim, vertex_count, simplex_count, V_coords, simplices, split_threshold,
                        far_theta * far_theta, near_theta * near_theta, intersection_theta * intersection_theta, ctx->stream );
        ev.stop();
        RP_HIP( hipStreamSynchronize( ctx->stream ) );
        finish_timing( ctx, 14 );
        if( counts )
        {
            counts[0] = B->C; counts[1] = B->L; counts[2] = B->depth; counts[3] = B->far_nnz; counts[4] = B->n_blocks;
            counts[5] = B->n_vf; counts[6] = B->n_inter; counts[7] = B->n_cols;
        }
        ctx->built = std::move( B );
    });
}

int rp_download_built_topology( rp_ctx * ctx, int32_t * P_ordering, int32_t * C_left, int32_t * C_right, int32_t * C_begin, int32_t * C_end,
                                int32_t * C_depth, double * C_serialized, int64_t * far_rp, int32_t * far_ci, int32_t * leaf_blocks,
                                int32_t * verynear_pairs, int32_t * inter_pairs, int64_t * col_ptr, int32_t * col_prim, uint32_t * col_mask,
                                int release )
{
    return guarded( "rp_download_built_topology", [&]()
    {
        require( ctx != nullptr && ctx->built, "no device-built topology (call rp_build_topology first)", RP_ENOTREADY );
        RP_HIP( hipSetDevice( ctx->device ) );
        BuiltTopology & B = *ctx->built;
        hipStream_t s = ctx->stream;
        auto get = [&]( auto * dst, const void * src, int64_t n )
        {
            if( dst && n > 0 ) { RP_HIP( hipMemcpyAsync( dst, src, (size_t)n * sizeof(*dst), hipMemcpyDeviceToHost, s ) ); }
        };
        get( P_ordering, B.final_ord, B.T );
        get( C_left, B.C_left.get(), B.C ); get( C_right, B.C_right.get(), B.C ); get( C_begin, B.C_begin.get(), B.C );
        get( C_end, B.C_end.get(), B.C ); get( C_depth, B.C_depth.get(), B.C ); get( C_serialized, B.C_ser.get(), 7 * B.C );
        get( far_rp, B.far_rp.get(), B.C + 1 ); get( far_ci, B.far_ci.get(), B.far_nnz );
        get( col_ptr, B.col_ptr.get(), B.L + 1 ); get( col_prim, B.col_prim.get(), B.n_cols ); get( col_mask, B.col_mask.get(), B.n_cols );
        // sorted (I << 32 | J) keys to (I, J) pairs
        std::vector<uint64_t> k;
        auto pairs = [&]( int32_t * dst, const uint64_t * src, int64_t n )
        {
            if( !dst || n <= 0 ) { return; }
            k.resize( (size_t)n );
            RP_HIP( hipMemcpyAsync( k.data(), src, (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToHost, s ) );
            RP_HIP( hipStreamSynchronize( s ) );
            for( int64_t i = 0; i < n; ++i ) { dst[2 * i] = (int32_t)( k[(size_t)i] >> 32 ); dst[2 * i + 1] = (int32_t)( k[(size_t)i] & 0xffffffffu ); }
        };
        pairs( leaf_blocks, B.blocks.get(), B.n_blocks );
        pairs( verynear_pairs, B.vf.get(), B.n_vf );
        pairs( inter_pairs, B.inter.get(), B.n_inter );
        RP_HIP( hipStreamSynchronize( s ) );
        if( release ) { ctx->built.reset(); }
    });
}

int rp_upload_data( rp_ctx * ctx, const double * V_coords, const double * P_near, const double * C_far )
{
    return guarded( "rp_upload_data", [&]()
    {
        require( ctx != nullptr && ctx->have_tree, "upload the tree first", RP_ENOTREADY );
        require( V_coords && P_near && C_far, "missing data arrays" );
        RP_HIP( hipSetDevice( ctx->device ) );
        const int64_t T = ctx->T, C = ctx->C, L = ctx->L;
        for( int64_t k = 0; k < 3 * ctx->V; ++k ) { require( std::isfinite( V_coords[k] ), "vertex coordinates contain NaN or Inf" ); }
        ctx->coords.upload( V_coords, (size_t)(3 * ctx->V), ctx->stream );
        ctx->stage.upload( P_near, (size_t)(ctx->nd * T), ctx->stream );
        ctx->p_a.require( (size_t)T ); ctx->p_x.require( (size_t)(3 * T) ); ctx->p_P.require( (size_t)(6 * T) ); ctx->p_rec.require( (size_t)(10 * T) );
        launch_prep_prims( T, ctx->nd, ctx->size, ctx->stage.get(), ctx->p_a.get(), ctx->p_x.get(), ctx->p_P.get(), ctx->p_rec.get(), ctx->stream );
        RP_HIP( hipStreamSynchronize( ctx->stream ) );
        ctx->stage.upload( C_far, (size_t)(10 * C), ctx->stream );
        ctx->c_a.require( (size_t)C ); ctx->c_x.require( (size_t)(3 * C) ); ctx->c_P.require( (size_t)(6 * C) );
        launch_prep_clusters( C, ctx->stage.get(), ctx->c_a.get(), ctx->c_x.get(), ctx->c_P.get(), ctx->stream );
        ctx->leaf_origin.require( (size_t)(3 * L) );
        launch_leaf_origin( L, C, ctx->leaf_cluster.get(), ctx->c_x.get(), ctx->leaf_origin.get(), ctx->stream );
        RP_HIP( hipGetLastError() );
        RP_HIP( hipStreamSynchronize( ctx->stream ) );
        ctx->have_data = true;
        ctx->have_state_rows = false;
        ctx->have_diff = ctx->have_metric = ctx->have_pl = ctx->have_h1 = false;
        ctx->have_vf_dh = ctx->have_vf_metric = ctx->have_pl_vf = false;
    });
}

int rp_update_coordinates( rp_ctx * ctx, const double * V_coords, int64_t ldV, int singletons, int ptr_flags )
{
    return guarded( "rp_update_coordinates", [&]()
    {
        require( ctx != nullptr && ctx->have_tree, "upload the mesh and the tree first", RP_ENOTREADY );
        require( V_coords != nullptr && ldV >= 3, "V_coords must be V x 3 with ldV >= 3" );
        RP_HIP( hipSetDevice( ctx->device ) );
        const int64_t T = ctx->T, C = ctx->C, L = ctx->L, V = ctx->V;
        hipStream_t s = ctx->stream;
        EventScope ev( ctx, 13 );
        if( ptr_flags & RP_DEVICE_PTR )
        {
            launch_copy3( V, V_coords, ldV, ctx->coords.get(), 3, s );
        }
        else
        {
            std::vector<double> h;
            pack_rows( V_coords, ldV, V, h );
            for( double x : h ) { require( std::isfinite( x ), "vertex coordinates contain NaN or Inf" ); }
            ctx->coords.upload( h.data(), h.size(), s );
        }
        ctx->stage.require( (size_t)( ctx->nd * T ) );
        ctx->upd_far.require( (size_t)( 10 * ( T + C ) ) );
        double * near = ctx->stage.get();
        double * pfar = ctx->upd_far.get();
        double * cfar = pfar + 10 * T;
        launch_near_far_rows( ctx->dom_dim, T, ctx->P_ordering.get(), ctx->simplices.get(), ctx->coords.get(), 3,
                              ctx->have_charges ? ctx->charges.get() : nullptr, near, pfar, s );
        launch_cluster_leaves( L, ctx->leaf_cluster.get(), ctx->C_begin.get(), ctx->C_end.get(), pfar, singletons, cfar, s );
        for( size_t d = 0; d + 1 < ctx->h_level_ptr.size(); ++d )
        {
            const int64_t b = ctx->h_level_ptr[d], e = ctx->h_level_ptr[d + 1];
            launch_cluster_level( e - b, ctx->level_nodes.get() + b, ctx->C_left.get(), ctx->C_right.get(), cfar, s );
        }
        ctx->p_a.require( (size_t)T ); ctx->p_x.require( (size_t)(3 * T) ); ctx->p_P.require( (size_t)(6 * T) ); ctx->p_rec.require( (size_t)(10 * T) );
        launch_prep_prims( T, ctx->nd, ctx->size, near, ctx->p_a.get(), ctx->p_x.get(), ctx->p_P.get(), ctx->p_rec.get(), s );
        ctx->c_a.require( (size_t)C ); ctx->c_x.require( (size_t)(3 * C) ); ctx->c_P.require( (size_t)(6 * C) );
        launch_prep_clusters( C, cfar, ctx->c_a.get(), ctx->c_x.get(), ctx->c_P.get(), s );
        ctx->leaf_origin.require( (size_t)(3 * L) );
        launch_leaf_origin( L, C, ctx->leaf_cluster.get(), ctx->c_x.get(), ctx->leaf_origin.get(), s );
        ev.stop();
        RP_HIP( hipGetLastError() );
        RP_HIP( hipStreamSynchronize( s ) );
        finish_timing( ctx, 13 );
        ctx->have_data = true;
        ctx->have_state_rows = true;
        ctx->have_diff = ctx->have_metric = ctx->have_pl = ctx->have_h1 = false;
        ctx->have_vf_dh = ctx->have_vf_metric = ctx->have_pl_vf = false;
    });
}

int rp_max_safe_step( rp_ctx * ctx, const double * U, int64_t ldU, double t_max, double TOL, int ptr_flags,
                      double * t_out, int64_t * stats )
{
    return guarded( "rp_max_safe_step", [&]()
    {
        require( ctx != nullptr && ctx->have_tree && ctx->have_data, "mesh, tree and data must be uploaded", RP_ENOTREADY );
        require( U != nullptr && ldU >= 3 && t_out != nullptr, "U must be V x 3 with ldU >= 3; t_out must not be NULL" );
        require( std::isfinite( t_max ) && t_max > 0.0, "max_time must be positive" );
        require( std::isfinite( TOL ) && TOL > 0.0 && TOL < 1.0, "TOL must lie in (0, 1)" );
        RP_HIP( hipSetDevice( ctx->device ) );
        const int64_t T = ctx->T, C = ctx->C, L = ctx->L, V = ctx->V;
        hipStream_t s = ctx->stream;
        ctx->col_U.require( (size_t)(3 * V) );
        if( ptr_flags & RP_DEVICE_PTR ) { launch_copy3( V, U, ldU, ctx->col_U.get(), 3, s ); }
        else
        {
            std::vector<double> h;
            pack_rows( U, ldU, V, h );
            for( double x : h ) { require( std::isfinite( x ), "update vectors contain NaN or Inf" ); }
            ctx->col_U.upload( h.data(), h.size(), s );
        }
        ctx->col_pbox.require( (size_t)(6 * T) ); ctx->col_cbox.require( (size_t)(6 * C) );
        ctx->col_SV.require( (size_t)( ctx->size * T ) ); ctx->col_PU.require( (size_t)( 6 * ctx->size * T ) );
        launch_prim_stage( ctx->dom_dim, T, ctx->P_ordering.get(), ctx->simplices.get(), ctx->coords.get(), ctx->col_U.get(), t_max,
                           ctx->col_pbox.get(), ctx->col_SV.get(), ctx->col_PU.get(), s );
        launch_leaf_boxes( L, ctx->leaf_cluster.get(), ctx->C_begin.get(), ctx->C_end.get(), ctx->col_pbox.get(), ctx->col_cbox.get(), s );
        for( size_t d = 0; d + 1 < ctx->h_level_ptr.size(); ++d )
        {
            const int64_t b = ctx->h_level_ptr[d], e = ctx->h_level_ptr[d + 1];
            launch_level_boxes( e - b, ctx->level_nodes.get() + b, ctx->C_left.get(), ctx->C_right.get(), ctx->col_cbox.get(), s );
        }
        // dual-tree traversal, one level per launch; work lists grow (and the traversal restarts) on overflow
        int64_t cap = std::max<int64_t>( (int64_t)1 << 20, 8 * C ), leaf_cap = std::max<int64_t>( (int64_t)1 << 20, 64 * L );
        ctx->col_counters.require( 4 );
        int64_t n_leaves = 0;
        for( int attempt = 0; ; ++attempt )
        {
            require( attempt < 12, "collision traversal work lists exceed 2^31 entries", RP_ENOMEM );
            ctx->col_list_a.require( (size_t)cap ); ctx->col_list_b.require( (size_t)cap ); ctx->col_leaves.require( (size_t)leaf_cap );
            int2 * in = ctx->col_list_a.get(), * out = ctx->col_list_b.get();
            const int2 root = make_int2( 0, 0 );
            RP_HIP( hipMemcpyAsync( in, &root, sizeof(int2), hipMemcpyHostToDevice, s ) );
            RP_HIP( hipMemsetAsync( ctx->col_counters.get(), 0, 3 * sizeof(unsigned long long), s ) );
            int64_t n_in = 1;
            bool overflow = false;
            while( n_in > 0 )
            {
                RP_HIP( hipMemsetAsync( ctx->col_counters.get(), 0, sizeof(unsigned long long), s ) );
                launch_traverse_step( n_in, in, out, cap, ctx->col_leaves.get(), leaf_cap, ctx->col_counters.get(),
                                      ctx->C_left.get(), ctx->C_right.get(), ctx->col_cbox.get(), s );
                unsigned long long cnt[3];
                RP_HIP( hipMemcpyAsync( cnt, ctx->col_counters.get(), sizeof(cnt), hipMemcpyDeviceToHost, s ) );
                RP_HIP( hipStreamSynchronize( s ) );
                if( cnt[2] ) { overflow = true; break; }
                n_in = (int64_t)cnt[0];
                n_leaves = (int64_t)cnt[1];
                std::swap( in, out );
            }
            if( !overflow ) { break; }
            cap *= 2; leaf_cap *= 2;
        }
        // candidate simplex pairs (non-adjacent, swept boxes meet), then the reference's bisection for each;
        // the candidate list grows to the count and is rebuilt when it overflows
        EventScope ev( ctx, 12 );
        int64_t n_cand = 0;
        for( int attempt = 0; attempt < 2; ++attempt )
        {
            ctx->col_cand.require( std::max<size_t>( ctx->col_cand.n, (size_t)std::max<int64_t>( (int64_t)1 << 22, 32 * n_leaves ) ) );
            const int64_t cand_cap = (int64_t)ctx->col_cand.n;
            RP_HIP( hipMemsetAsync( ctx->col_counters.get() + 3, 0, sizeof(unsigned long long), s ) );
            launch_leaf_pair_candidates( ctx->dom_dim, n_leaves, ctx->col_leaves.get(), ctx->C_begin.get(), ctx->C_end.get(), ctx->col_SV.get(),
                                         ctx->col_pbox.get(), ctx->col_cbox.get(), ctx->col_cand.get(), cand_cap, ctx->col_counters.get() + 3, s );
            unsigned long long cnt = 0;
            RP_HIP( hipMemcpyAsync( &cnt, ctx->col_counters.get() + 3, sizeof(cnt), hipMemcpyDeviceToHost, s ) );
            RP_HIP( hipStreamSynchronize( s ) );
            n_cand = (int64_t)cnt;
            if( n_cand <= cand_cap ) { break; }
            require( attempt == 0, "collision candidate list overflow", RP_ENOMEM );
            ctx->col_cand.require( (size_t)n_cand );
        }
        unsigned long long res[3];
        double tm = t_max;
        std::memcpy( &res[0], &tm, sizeof(double) );
        res[1] = (unsigned long long)n_cand; res[2] = 0;
        RP_HIP( hipMemcpyAsync( ctx->col_counters.get(), res, sizeof(res), hipMemcpyHostToDevice, s ) );
        launch_pair_steps( ctx->dom_dim, n_cand, ctx->col_cand.get(), ctx->col_PU.get(), t_max, TOL, ctx->col_counters.get(), s );
        ev.stop();
        RP_HIP( hipGetLastError() );
        RP_HIP( hipMemcpyAsync( res, ctx->col_counters.get(), sizeof(res), hipMemcpyDeviceToHost, s ) );
        RP_HIP( hipStreamSynchronize( s ) );
        finish_timing( ctx, 12 );
        double t;
        std::memcpy( &t, &res[0], sizeof(double) );
        *t_out = t;
        if( stats ) { stats[0] = n_leaves; stats[1] = (int64_t)res[1]; stats[2] = (int64_t)res[2]; }
    });
}

int rp_download_state_rows( rp_ctx * ctx, double * P_near, double * C_far, double * V_coords )
{
    return guarded( "rp_download_state_rows", [&]()
    {
        require( ctx != nullptr && ctx->have_state_rows, "no rows from rp_update_coordinates", RP_ENOTREADY );
        RP_HIP( hipSetDevice( ctx->device ) );
        if( P_near ) { ctx->stage.download( P_near, (size_t)( ctx->nd * ctx->T ), ctx->stream ); }
        if( C_far ) { RP_HIP( hipMemcpyAsync( C_far, ctx->upd_far.get() + 10 * ctx->T, sizeof(double) * 10 * ctx->C, hipMemcpyDeviceToHost, ctx->stream ) ); }
        if( V_coords ) { ctx->coords.download( V_coords, (size_t)( 3 * ctx->V ), ctx->stream ); }
        RP_HIP( hipStreamSynchronize( ctx->stream ) );
    });
}

int rp_tp0_compute( rp_ctx * ctx, unsigned flags, double q, double p, int precision, double * energy_out )
{
    return guarded( "rp_tp0_compute", [&]()
    {
        require( ctx != nullptr && ctx->have_data && ctx->have_near && ctx->have_far, "mesh, tree, patterns and data must be uploaded", RP_ENOTREADY );
        require( (flags & (RP_ENERGY | RP_DIFF | RP_METRIC)) != 0, "no work requested" );
        require( precision == RP_FP64 || precision == RP_MIXED || precision == RP_MIXED_REFINED, "unknown precision" );
        require( q >= 0, "Tangent-point kernels not well-defined for q < 0." );
        require( q == 0 || q >= 2, "Tangent-point kernels not implement for 0 < q < 2." );
        RP_HIP( hipSetDevice( ctx->device ) );
        const int n = ctx->dom_dim;
        const int64_t T = ctx->T, C = ctx->C, L = ctx->L;
        const double s = (p - n) / q;
        const double s_exp = -0.5 * ( 2.0 * (s - 1.0) + n );
        const Expo ex = make_expo( q, p, n, s_exp );
        const int kind = exponent_kind( q, p );
        const bool mixed = ( precision == RP_MIXED );
        const bool E = flags & RP_ENERGY, D = flags & RP_DIFF, M = flags & RP_METRIC;
        const bool copies = M && precision == RP_MIXED_REFINED;

        const bool have_vf = ctx->vf_nnz > 0;
        const int64_t R = have_vf ? ctx->vf_nrows : 0;
        ctx->partials.require( (size_t)(L + C + R) );
        ctx->scalar.require( (size_t)(1 + REDUCE_BLOCKS + 3) );
        if( D ) { ctx->nf_dx.require( (size_t)(10 * T) ); ctx->ff_dx.require( (size_t)(10 * C) ); }
        if( M )
        {
            ctx->nf_vals.require( (size_t)std::max<int64_t>( 1, 2 * ctx->nf_slots ) );
            ctx->nf_diag.require( (size_t)(2 * T) );
            if( copies )
            {
                ctx->nf_vals32.require( (size_t)std::max<int64_t>( 1, ctx->nf_slots ) );
                ctx->nf_diag32.require( (size_t)(2 * T) );
            }
            ctx->ff_vals.require( (size_t)std::max<int64_t>( 1, 2 * ctx->far_nnz ) );
            ctx->ff_diag.require( (size_t)(2 * C) );
        }
        EventScope ev( ctx, 0 );
        if( E ) { launch_fill( ctx->partials.get(), L + C, 0.0, ctx->stream ); }
        const bool have_nf = ctx->nf_cols > 0;
        const bool have_ff = ctx->far_nnz > 0;
        if( !have_nf )
        {
            if( D ) { launch_fill( ctx->nf_dx.get(), 10 * T, 0.0, ctx->stream ); }
            if( M ) { launch_fill( ctx->nf_diag.get(), 2 * T, 0.0, ctx->stream ); }
            if( copies ) { launch_fill( ctx->nf_diag32.get(), 2 * T, 0.0, ctx->stream ); }
        }
        if( !have_ff )
        {
            if( D ) { launch_fill( ctx->ff_dx.get(), 10 * C, 0.0, ctx->stream ); }
            if( M ) { launch_fill( ctx->ff_diag.get(), 2 * C, 0.0, ctx->stream ); }
        }
        RP_HIP( hipEventRecord( ctx->ev[4][0], ctx->stream ) );
        if( have_nf )
        {
            NFArgs A{};
            A.L = L; A.T = T;
            A.leaf_begin = ctx->leaf_begin.get(); A.leaf_end = ctx->leaf_end.get();
            A.col_ptr = ctx->col_ptr.get(); A.col_prim = ctx->col_prim.get(); A.col_mask = ctx->col_mask.get(); A.val_ptr = ctx->val_ptr.get();
            A.a = ctx->p_a.get(); A.x = ctx->p_x.get(); A.P = ctx->p_P.get(); A.rec = ctx->p_rec.get(); A.origin = ctx->leaf_origin.get();
            A.scale = 1.0 / ctx->size; A.size = ctx->size; A.ex = ex;
            A.dx = D ? ctx->nf_dx.get() : nullptr;
            A.epart = ctx->partials.get();
            A.vals = M ? (void *)ctx->nf_vals.get() : nullptr;
            A.diag = M ? ctx->nf_diag.get() : nullptr;
            A.vals32 = copies ? reinterpret_cast<float2 *>( ctx->nf_vals32.get() ) : nullptr;
            A.diag32 = copies ? ctx->nf_diag32.get() : nullptr;
            launch_nf( flags, kind, mixed, A, ctx->stream );
            RP_HIP( hipGetLastError() );
        }
        RP_HIP( hipEventRecord( ctx->ev[4][1], ctx->stream ) );
        RP_HIP( hipEventRecord( ctx->ev[5][0], ctx->stream ) );
        if( have_ff )
        {
            FFArgs B{};
            B.C = C; B.rp = ctx->far_rp.get(); B.ci = ctx->far_ci.get();
            B.a = ctx->c_a.get(); B.x = ctx->c_x.get(); B.P = ctx->c_P.get(); B.ex = ex;
            B.dx = D ? ctx->ff_dx.get() : nullptr;
            B.epart = ctx->partials.get() + L;
            B.vals = M ? (void *)ctx->ff_vals.get() : nullptr;
            B.diag = M ? ctx->ff_diag.get() : nullptr;
            launch_ff( flags, kind, mixed, B, ctx->stream );
            RP_HIP( hipGetLastError() );
        }
        RP_HIP( hipEventRecord( ctx->ev[5][1], ctx->stream ) );
        RP_HIP( hipEventRecord( ctx->ev[11][0], ctx->stream ) );
        double * vf_stat = ctx->scalar.get() + 1 + REDUCE_BLOCKS;
        if( D ) { ctx->have_vf_dh = false; }
        if( M ) { ctx->have_vf_metric = false; }
        if( have_vf )
        {
            // very-near field in FP64: per-entry adaptive traversals, then per-row sums
            const int64_t nnz = ctx->vf_nnz;
            ctx->vf_rec.require( (size_t)(VF_REC * nnz) );
            if( D ) { ctx->vf_dh.require( (size_t)(3 * ctx->size * T) ); launch_fill( ctx->vf_dh.get(), 3 * ctx->size * T, 0.0, ctx->stream ); }
            if( M )
            {
                ctx->vf_vals.require( (size_t)(2 * nnz) ); ctx->vf_diag.require( (size_t)(2 * T) );
                launch_fill( ctx->vf_diag.get(), 2 * T, 0.0, ctx->stream );
            }
            VFArgs A{};
            A.nnz = nnz; A.T = T; A.row = ctx->vf_row.get(); A.col = ctx->vf_ci.get();
            A.P_ordering = ctx->P_ordering.get(); A.simplices = ctx->simplices.get(); A.coords = ctx->coords.get();
            A.a = ctx->p_a.get(); A.P = ctx->p_P.get();
            A.theta2 = ctx->vf_theta * ctx->vf_theta; A.max_ref = ctx->vf_max_ref; A.ex = ex;
            A.energy = E; A.diff = D; A.metric = M; A.pl_sexp = 0.0;
            A.cap = VF_LANE_EVALUATION_LIMIT; A.rec = ctx->vf_rec.get();
            launch_vf_pairs( n, 0, A, ctx->stream );
            VFRowArgs W{};
            W.nrows = R; W.T = T; W.rows = ctx->vf_rows.get(); W.rp = ctx->vf_rp.get(); W.dg = ctx->vf_dg.get(); W.tr = ctx->vf_tr.get();
            W.rec = ctx->vf_rec.get();
            W.dx = D ? ctx->nf_dx.get() : nullptr; W.dh = D ? ctx->vf_dh.get() : nullptr;
            W.vals = M ? ctx->vf_vals.get() : nullptr; W.diag = M ? ctx->vf_diag.get() : nullptr;
            W.epart = E ? ctx->partials.get() + L + C : nullptr;
            launch_vf_rows( n, 0, W, ctx->stream );
            launch_vf_stats( nnz, ctx->vf_rec.get(), vf_stat, ctx->stream );
            RP_HIP( hipGetLastError() );
        }
        RP_HIP( hipEventRecord( ctx->ev[11][1], ctx->stream ) );
        if( E ) { launch_reduce_sum( ctx->partials.get(), L + C + R, ctx->scalar.get(), ctx->stream ); }
        ev.stop();
        RP_HIP( hipGetLastError() );
        double en = 0;
        if( E ) { RP_HIP( hipMemcpyAsync( &en, ctx->scalar.get(), sizeof(double), hipMemcpyDeviceToHost, ctx->stream ) ); }
        double st[3] = { 0, 0, 0 };
        if( have_vf ) { RP_HIP( hipMemcpyAsync( st, vf_stat, sizeof(st), hipMemcpyDeviceToHost, ctx->stream ) ); }
        RP_HIP( hipStreamSynchronize( ctx->stream ) );
        finish_timing( ctx, 0 ); finish_timing( ctx, 4 ); finish_timing( ctx, 5 ); finish_timing( ctx, 11 );
        for( int k = 0; k < 3; ++k ) { ctx->vf_stats[k] = st[k]; }
        if( st[2] != 0.0 ) { ctx->have_diff = ctx->have_metric = false; }
        require( st[2] == 0.0, "very-near-field subdivision exceeded 2^24 evaluations on one level-1 sub-pair "
                 "(primitives almost touching); refine the mesh or lower adaptivity_settings.max_refinement", RP_EUNSUPPORTED );
        if( energy_out ) { *energy_out = E ? en : 0.0; }
        if( D ) { ctx->have_diff = true; ctx->have_vf_dh = have_vf; }
        if( M ) { ctx->have_metric = true; ctx->have_vf_metric = have_vf; ctx->metric_precision = precision; ++ctx->metric_version; }
    });
}

int rp_tp0_energy_noalloc( rp_ctx * ctx, double q, double p, int precision, double * energy_out, int ptr_flags )
{
    return guarded( "rp_tp0_energy_noalloc", [&]()
    {
        require( ctx != nullptr && ctx->have_data && ctx->have_near && ctx->have_far, "mesh, tree, patterns and data must be uploaded", RP_ENOTREADY );
        require( precision == RP_FP64 || precision == RP_MIXED || precision == RP_MIXED_REFINED, "unknown precision" );
        require( q >= 0, "Tangent-point kernels not well-defined for q < 0." );
        require( q == 0 || q >= 2, "Tangent-point kernels not implement for 0 < q < 2." );
        require( energy_out != nullptr, "energy_out is NULL" );
        RP_HIP( hipSetDevice( ctx->device ) );
        double * base = scratch_address( ctx->device );
        require( base != nullptr, "no scratch array on this device (device index beyond 63)", RP_EUNSUPPORTED );
        const int n = ctx->dom_dim;
        const int64_t T = ctx->T, C = ctx->C, L = ctx->L;
        const bool have_vf = ctx->vf_nnz > 0;
        const int64_t R = have_vf ? ctx->vf_nrows : 0;
        const int64_t nnz = have_vf ? ctx->vf_nnz : 0;
        // scratch layout, each part starting on a 256-byte boundary: energy partials (leaves, clusters,
        // very-near rows), the reduction (1 + REDUCE_BLOCKS) and the very-near statistics (3), the
        // very-near per-entry records (VF_REC per entry)
        auto up32 = []( int64_t x ) { return ( x + 31 ) / 32 * 32; };
        const int64_t o_scalar = up32( L + C + R );
        const int64_t o_rec = o_scalar + up32( 1 + REDUCE_BLOCKS + 3 );
        const int64_t need = o_rec + VF_REC * nnz;
        require( need <= RP_NOALLOC_SCRATCH_DOUBLES, "mesh needs " + std::to_string( need ) + " doubles of scratch; rp_tp0_energy_noalloc holds at most " +
                 std::to_string( (long long)RP_NOALLOC_SCRATCH_DOUBLES ) + " (use rp_tp0_compute)" );
        double * partials = base;
        double * scalar = base + o_scalar;
        double * vf_stat = scalar + 1 + REDUCE_BLOCKS;
        double * rec = base + o_rec;
        const double s = (p - n) / q;
        const double s_exp = -0.5 * ( 2.0 * (s - 1.0) + n );
        const Expo ex = make_expo( q, p, n, s_exp );
        const int kind = exponent_kind( q, p );
        const bool mixed = ( precision == RP_MIXED );
        // the array keeps the last call's contents: the partials start at zero
        launch_fill( partials, L + C, 0.0, ctx->stream );
        if( ctx->nf_cols > 0 )
        {
            NFArgs A{};
            A.L = L; A.T = T;
            A.leaf_begin = ctx->leaf_begin.get(); A.leaf_end = ctx->leaf_end.get();
            A.col_ptr = ctx->col_ptr.get(); A.col_prim = ctx->col_prim.get(); A.col_mask = ctx->col_mask.get(); A.val_ptr = ctx->val_ptr.get();
            A.a = ctx->p_a.get(); A.x = ctx->p_x.get(); A.P = ctx->p_P.get(); A.rec = ctx->p_rec.get(); A.origin = ctx->leaf_origin.get();
            A.scale = 1.0 / ctx->size; A.size = ctx->size; A.ex = ex;
            A.epart = partials;
            launch_nf( RP_ENERGY, kind, mixed, A, ctx->stream );
            RP_HIP( hipGetLastError() );
        }
        if( ctx->far_nnz > 0 )
        {
            FFArgs B{};
            B.C = C; B.rp = ctx->far_rp.get(); B.ci = ctx->far_ci.get();
            B.a = ctx->c_a.get(); B.x = ctx->c_x.get(); B.P = ctx->c_P.get(); B.ex = ex;
            B.epart = partials + L;
            launch_ff( RP_ENERGY, kind, mixed, B, ctx->stream );
            RP_HIP( hipGetLastError() );
        }
        if( have_vf )
        {
            VFArgs A{};
            A.nnz = nnz; A.T = T; A.row = ctx->vf_row.get(); A.col = ctx->vf_ci.get();
            A.P_ordering = ctx->P_ordering.get(); A.simplices = ctx->simplices.get(); A.coords = ctx->coords.get();
            A.a = ctx->p_a.get(); A.P = ctx->p_P.get();
            A.theta2 = ctx->vf_theta * ctx->vf_theta; A.max_ref = ctx->vf_max_ref; A.ex = ex;
            A.energy = true; A.diff = false; A.metric = false; A.pl_sexp = 0.0;
            A.cap = VF_LANE_EVALUATION_LIMIT; A.rec = rec;
            launch_vf_pairs( n, 0, A, ctx->stream );
            VFRowArgs W{};
            W.nrows = R; W.T = T; W.rows = ctx->vf_rows.get(); W.rp = ctx->vf_rp.get(); W.dg = ctx->vf_dg.get(); W.tr = ctx->vf_tr.get();
            W.rec = rec;
            W.epart = partials + L + C;
            launch_vf_rows( n, 0, W, ctx->stream );
            launch_vf_stats( nnz, rec, vf_stat, ctx->stream );
            RP_HIP( hipGetLastError() );
        }
        launch_reduce_sum( partials, L + C + R, scalar, ctx->stream );
        RP_HIP( hipGetLastError() );
        const hipMemcpyKind dir = ( ptr_flags & RP_DEVICE_PTR ) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
        RP_HIP( hipMemcpyAsync( energy_out, scalar, sizeof(double), dir, ctx->stream ) );
        double st[3] = { 0, 0, 0 };
        if( have_vf ) { RP_HIP( hipMemcpyAsync( st, vf_stat, sizeof(st), hipMemcpyDeviceToHost, ctx->stream ) ); }
        RP_HIP( hipStreamSynchronize( ctx->stream ) );
        require( st[2] == 0.0, "very-near-field subdivision exceeded 2^24 evaluations on one level-1 sub-pair "
                 "(primitives almost touching); refine the mesh or lower adaptivity_settings.max_refinement", RP_EUNSUPPORTED );
    });
}

int rp_download_derivatives( rp_ctx * ctx, double * P_D_near, double * P_D_far )
{
    return guarded( "rp_download_derivatives", [&]()
    {
        require( ctx != nullptr && ctx->have_diff, "no differential computed", RP_ENOTREADY );
        RP_HIP( hipSetDevice( ctx->device ) );
        const int64_t T = ctx->T, L = ctx->L;
        collect_far( ctx );
        std::vector<double> dx( (size_t)(10 * T) ), lf( (size_t)(10 * L) );
        std::vector<int32_t> ord( (size_t)T ), lop( (size_t)T );
        ctx->nf_dx.download( dx.data(), dx.size(), ctx->stream );
        ctx->leaf_far.download( lf.data(), lf.size(), ctx->stream );
        ctx->P_ordering.download( ord.data(), ord.size(), ctx->stream );
        ctx->leaf_of_prim.download( lop.data(), lop.size(), ctx->stream );
        std::vector<double> dh;
        if( ctx->have_vf_dh ) { dh.resize( (size_t)(3 * ctx->size * T) ); ctx->vf_dh.download( dh.data(), dh.size(), ctx->stream ); }
        RP_HIP( hipStreamSynchronize( ctx->stream ) );
        const int nd = ctx->nd, size = ctx->size;
        for( int64_t i = 0; i < T; ++i )
        {
            const int64_t t = ord[(size_t)i];
            if( P_D_near )
            {
                double * o = P_D_near + nd * t;
                o[0] = dx[(size_t)i];
                for( int h = 0; h < size; ++h ) { for( int c = 0; c < 3; ++c ) { o[1 + 3 * h + c] = dx[(size_t)((1 + c) * T + i)]; } }
                if( !dh.empty() ) { for( int h = 0; h < 3 * size; ++h ) { o[1 + h] += dh[(size_t)(3 * size * i + h)]; } }
                for( int k = 0; k < 6; ++k ) { o[1 + 3 * size + k] = dx[(size_t)((4 + k) * T + i)]; }
            }
            if( P_D_far )
            {
                const int64_t l = lop[(size_t)i];
                for( int k = 0; k < 10; ++k ) { P_D_far[10 * t + k] = lf[(size_t)(k * L + l)]; }
            }
        }
    });
}

int rp_assemble_vertex_gradient( rp_ctx * ctx, double alpha, double beta, double * Y, int64_t ldY, int ptr_flags )
{
    return guarded( "rp_assemble_vertex_gradient", [&]()
    {
        require( ctx != nullptr && ctx->have_diff, "no differential computed", RP_ENOTREADY );
        require( Y != nullptr && ldY >= 3, "Y must be V x 3 with ldY >= 3" );
        RP_HIP( hipSetDevice( ctx->device ) );
        const int64_t T = ctx->T, V = ctx->V, L = ctx->L;
        const bool dev = ptr_flags & RP_DEVICE_PTR;
        std::vector<double> h, hy;
        double * Yd = Y;
        int64_t ld = ldY;
        if( !dev )
        {
            ctx->Yd.require( (size_t)(3 * V) );
            if( beta != 0.0 ) { pack_rows( Y, ldY, V, h ); ctx->Yd.upload( h.data(), h.size(), ctx->stream ); }
            Yd = ctx->Yd.get(); ld = 3;
        }
        ctx->hull_grad.require( (size_t)(T * ctx->size * 3) );
        EventScope ev( ctx, 1 );
        collect_far( ctx );
        launch_hulls( ctx->dom_dim, T, L, ctx->P_ordering.get(), ctx->leaf_of_prim.get(), ctx->simplices.get(), ctx->coords.get(),
                      ctx->have_charges ? ctx->charges.get() : nullptr, ctx->nf_dx.get(), ctx->have_vf_dh ? ctx->vf_dh.get() : nullptr,
                      ctx->leaf_far.get(), ctx->hull_grad.get(), ctx->stream );
        launch_assemble( V, ctx->vc_ptr.get(), ctx->vc_corner.get(), ctx->hull_grad.get(), alpha, beta, Yd, ld, ctx->stream );
        ev.stop();
        RP_HIP( hipGetLastError() );
        if( !dev )
        {
            hy.resize( (size_t)(3 * V) );
            ctx->Yd.download( hy.data(), hy.size(), ctx->stream );
        }
        RP_HIP( hipStreamSynchronize( ctx->stream ) );
        finish_timing( ctx, 1 );
        if( !dev ) { for( int64_t v = 0; v < V; ++v ) { for( int k = 0; k < 3; ++k ) { Y[ldY * v + k] = hy[(size_t)(3 * v + k)]; } } }
    });
}

int rp_download_vf_metric( rp_ctx * ctx, double * vf_vals, double * vf_diag )
{
    return guarded( "rp_download_vf_metric", [&]()
    {
        require( ctx != nullptr && ctx->have_metric, "no metric computed", RP_ENOTREADY );
        RP_HIP( hipSetDevice( ctx->device ) );
        const int64_t T = ctx->T;
        if( !ctx->have_vf_metric )
        {
            if( vf_diag ) { std::fill( vf_diag, vf_diag + 2 * T, 0.0 ); }
            return;
        }
        if( vf_vals ) { ctx->vf_vals.download( vf_vals, (size_t)(2 * ctx->vf_nnz), ctx->stream ); }
        if( vf_diag ) { ctx->vf_diag.download( vf_diag, (size_t)(2 * T), ctx->stream ); }
        RP_HIP( hipStreamSynchronize( ctx->stream ) );
    });
}

int rp_vf_stats( rp_ctx * ctx, int64_t * evaluations, int * max_level )
{
    return guarded( "rp_vf_stats", [&]()
    {
        require( ctx != nullptr, "ctx is NULL" );
        if( evaluations ) { *evaluations = (int64_t)ctx->vf_stats[0]; }
        if( max_level ) { *max_level = (int)ctx->vf_stats[1]; }
    });
}

int rp_download_metric( rp_ctx * ctx, double * nf_vals, double * nf_diag, double * ff_vals, double * ff_diag )
{
    return guarded( "rp_download_metric", [&]()
    {
        require( ctx != nullptr && ctx->have_metric, "no metric computed", RP_ENOTREADY );
        RP_HIP( hipSetDevice( ctx->device ) );
        const bool mixed = ctx->metric_precision == RP_MIXED;
        auto get_vals = [&]( const DBuf<double> & buf, int64_t pairs, double * out )
        {
            if( !out || pairs == 0 ) { return; }
            if( !mixed ) { buf.download( out, (size_t)(2 * pairs), ctx->stream ); return; }
            std::vector<float> f( (size_t)(2 * pairs) );
            RP_HIP( hipMemcpyAsync( f.data(), buf.get(), f.size() * sizeof(float), hipMemcpyDeviceToHost, ctx->stream ) );
            RP_HIP( hipStreamSynchronize( ctx->stream ) );
            for( size_t k = 0; k < f.size(); ++k ) { out[k] = f[k]; }
        };
        get_vals( ctx->nf_vals, ctx->nf_cols > 0 ? ctx->nf_slots : 0, nf_vals );
        get_vals( ctx->ff_vals, ctx->far_nnz, ff_vals );
        if( nf_diag ) { ctx->nf_diag.download( nf_diag, (size_t)(2 * ctx->T), ctx->stream ); }
        if( ff_diag ) { ctx->ff_diag.download( ff_diag, (size_t)(2 * ctx->C), ctx->stream ); }
        RP_HIP( hipStreamSynchronize( ctx->stream ) );
    });
}

int rp_multiply_metric( rp_ctx * ctx, double alpha, const double * X, int64_t ldX,
                        double beta, double * Y, int64_t ldY, int nrhs,
                        int vf_flag, int nf_flag, int ff_flag, int ptr_flags )
{
    return guarded( "rp_multiply_metric", [&]()
    {
        require( ctx != nullptr, "ctx is NULL" );
        require( nrhs == 3, "nrhs = " + std::to_string(nrhs) + " != 3 = AMB_DIM. The current implementation only accepts nrhs = AMB_DIM. Doing nothing." );
        require( ctx->have_metric, "metric values not assembled (call rp_tp0_compute with RP_METRIC)", RP_ENOTREADY );
        require( X != nullptr && Y != nullptr && ldX >= 3 && ldY >= 3, "X and Y must be V x 3 with ld >= 3" );
        require( nf_flag != RP_NF_FP32_COPIES || ctx->metric_precision == RP_MIXED_REFINED,
                 "RP_NF_FP32_COPIES needs a metric assembled at RP_MIXED_REFINED", RP_ENOTREADY );
        (void)vf_flag;
        RP_HIP( hipSetDevice( ctx->device ) );
        const int64_t V = ctx->V;
        const bool dev = ptr_flags & RP_DEVICE_PTR;
        std::vector<double> h, hx, hy;
        const double * Xd = X;
        int64_t ldx = ldX;
        double * Yd = Y;
        int64_t ldy = ldY;
        if( !dev )
        {
            pack_rows( X, ldX, V, hx );
            ctx->Xd.upload( hx.data(), hx.size(), ctx->stream );
            Xd = ctx->Xd.get(); ldx = 3;
            ctx->Yd.require( (size_t)(3 * V) );
            if( beta != 0.0 ) { pack_rows( Y, ldY, V, h ); ctx->Yd.upload( h.data(), h.size(), ctx->stream ); }
            Yd = ctx->Yd.get(); ldy = 3;
        }
        EventScope ev( ctx, 2 );
        metric_apply( ctx, alpha, Xd, ldx, beta, Yd, ldy, nf_flag != 0, ff_flag != 0, true, vf_flag != 0, nf_flag == RP_NF_FP32_COPIES );
        ev.stop();
        RP_HIP( hipGetLastError() );
        if( !dev )
        {
            hy.resize( (size_t)(3 * V) );
            ctx->Yd.download( hy.data(), hy.size(), ctx->stream );
        }
        RP_HIP( hipStreamSynchronize( ctx->stream ) );
        finish_timing( ctx, 2 ); finish_timing( ctx, 6 );
        if( !dev ) { for( int64_t v = 0; v < V; ++v ) { for( int k = 0; k < 3; ++k ) { Y[ldY * v + k] = hy[(size_t)(3 * v + k)]; } } }
    });
}

int rp_last_timing( rp_ctx * ctx, int which, double * ms )
{
    return guarded( "rp_last_timing", [&]()
    {
        require( ctx != nullptr && ms != nullptr && which >= 0 && which < 16, "bad arguments" );
        *ms = ctx->timings.ms[which];
    });
}

} // extern "C"
