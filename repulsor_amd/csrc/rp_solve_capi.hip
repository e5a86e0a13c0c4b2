// C-ABI entry points of the metric solve (declared in include/repulsor_hip.h): the pseudo-Laplacian
// L_{s'}, the H^1 factor and its triangular solves, the preconditioner P0 = H^-1 L_{2-s} H^-1 and the
// preconditioned conjugate-gradient solve of the tangent-point metric.

#include "rp_capi_util.hpp"
#include "k_pair.hpp"
#include "rp_launch.hpp"

#include <algorithm>
#include <cstring>

namespace rp
{
    // Y = alpha L X + beta Y (device vectors). Reference: MetricDimRestricted::MultiplyMetric with the
    // LowOrder pre/post (src/Metrics/MetricDimRestricted.hpp:109-139) around
    // PseudoLaplacian_Traversor::MultiplyMetric (src/Metrics/PseudoLaplacian/PseudoLaplacian_Traversor.hpp:143-300).
    // xcols (may be null): the pre-operator's vertex index per primitive slot in place of pre_cols (X in
    // another numbering); yperm (may be null): row j of Y takes vertex yperm[j].
    static void pl_apply( rp_ctx * ctx, double alpha, const double * Xd, int64_t ldx, double beta, double * Yd, int64_t ldy, bool nf_flag, bool ff_flag,
                          const int32_t * xcols = nullptr, const int32_t * yperm = nullptr )
    {
        const int64_t T = ctx->T, V = ctx->V, C = ctx->C, L = ctx->L;
        const bool mixed = ctx->pl_precision != RP_FP64;   // the far field: FP32 unless FP64
        ctx->P_in.require( (size_t)(12 * T) ); ctx->P_out.require( (size_t)(12 * T) );
        ctx->C_in.require( (size_t)(12 * C) ); ctx->C_out.require( (size_t)(12 * C) );
        ctx->leaf_out.require( (size_t)(12 * L) );
        const bool do_nf = nf_flag && ctx->nf_cols > 0;
        const bool do_ff = ff_flag && ctx->far_nnz > 0;
        launch_pl_pre( ctx->size, T, xcols ? xcols : ctx->pre_cols.get(), ctx->pre_vals.get(), Xd, ldx, ctx->P_in.get(), ctx->stream );
        if( do_ff )
        {
            launch_up_leaves( L, ctx->leaf_cluster.get(), ctx->C_begin.get(), ctx->C_end.get(), ctx->P_in.get(), ctx->C_in.get(), ctx->stream, 3 );
            launch_up_tree( ctx->h_level_ptr, ctx->level_nodes.get(), ctx->C_left.get(), ctx->C_right.get(), ctx->C_in.get(), ctx->stream, 3 );
            launch_pl_ff_spmm( mixed, C, ctx->far_rp.get(), ctx->far_ci.get(), ctx->pl_ff_vals.get(), ctx->pl_ff_diag.get(), ctx->C_in.get(), ctx->C_out.get(), ctx->stream );
            launch_leaf_out3( L, C, ctx->leaf_anc.get(), ctx->leaf_anc_len.get(), ctx->C_out.get(), ctx->leaf_out.get(), ctx->stream );
        }
        double * S_out = ctx->P_out.get();   // T x (n+1) x 3
        if( do_nf )
        {
            launch_pl_nf_spmm( ctx->pl_precision, ctx->size, L, ctx->leaf_begin.get(), ctx->leaf_end.get(), ctx->col_ptr.get(), ctx->col_prim.get(),
                               ctx->col_mask.get(), ctx->val_ptr.get(), ctx->pl_nf_vals.get(), ctx->pl_nf_diag.get(), ctx->P_in.get(),
                               ctx->leaf_out.get(), do_ff ? 1 : 0, ctx->pre_vals.get(), S_out, ctx->stream );
        }
        else
        {
            launch_pl_post_prim( ctx->size, T, ctx->leaf_of_prim.get(), ctx->pre_vals.get(), ctx->leaf_out.get(), do_ff ? 1 : 0, S_out, ctx->stream );
        }
        if( ctx->have_pl_vf )
        {
            launch_vf_pl_spmm_post( ctx->size, ctx->vf_nrows, ctx->vf_rows.get(), ctx->vf_rp.get(), ctx->vf_ci.get(), ctx->pl_vf_vals.get(),
                                    ctx->pl_vf_diag.get(), ctx->P_in.get(), ctx->pre_vals.get(), S_out, ctx->stream );
        }
        launch_post_vert( V, ctx->vp_ptr.get(), ctx->vp_entry.get(), S_out, alpha, beta, Yd, ldy, ctx->stream, yperm );
    }

    static CholDev chol_dev( rp_ctx * ctx )
    {
        CholDev F{};
        F.sn_begin = ctx->h1_sn_begin.get(); F.row_ptr = ctx->h1_row_ptr.get(); F.rows = ctx->h1_rows.get();
        F.val_ptr = ctx->h1_val_ptr.get(); F.vals = ctx->h1_vals.get(); F.vals_t = ctx->h1_vals_t.get();
        F.kid_ptr = ctx->h1_kid_ptr.get(); F.kids = ctx->h1_kids.get(); F.cmap = ctx->h1_cmap.get();
        F.perm = ctx->h1_perm.get(); F.z = ctx->h1_z.get(); F.xp = ctx->h1_xp.get();
        F.pull2 = reinterpret_cast<const int2 *>( ctx->h1_pull2.get() ); F.pull_src = ctx->h1_pull_src.get(); F.bp = ctx->h1_bp.get();
        return F;
    }

    // X = alpha H^-1 B + beta X (device vectors, V x 3): one launch per elimination level and sweep.
    // Bd null: the right-hand side is already in h1_bp (alpha B in the factor's numbering); Xd null: the
    // solution stays in h1_xp (factor numbering).
    static void h1_solve( rp_ctx * ctx, double alpha, const double * Bd, int64_t ldb, double beta, double * Xd, int64_t ldx )
    {
        const CholDev F = chol_dev( ctx );
        const size_t nl = ctx->h1_level_small.size();
        const CholTask * tasks = ctx->h1_tasks.get();
        if( Bd ) { launch_chol_gather( ctx->h1_n, ctx->h1_perm.get(), alpha, Bd, ldb, ctx->h1_bp.get(), ctx->stream ); }
        for( size_t l = 0; l < nl; ++l )
        {
            if( ctx->h1_level_small[l] )
            {
                const int64_t b = ctx->h1_level_ptr[l], e = ctx->h1_level_ptr[l + 1];
                launch_chol_fwd_small( F, ctx->h1_sdesc.get() + b, e - b, ctx->h1_level_g[l], ctx->h1_level_rmax[l], ctx->stream );
            }
            else if( ctx->h1_level_split[l] )
            {
                const int64_t tb = ctx->h1_task_ptr[0][l], te = ctx->h1_task_ptr[0][l + 1];
                const int64_t sb = ctx->h1_task_ptr[3][l], se = ctx->h1_task_ptr[3][l + 1];
                launch_chol_fwd_split( F, tasks + tb, te - tb, tasks + sb, se - sb, ctx->h1_fr.get(), ctx->stream );
            }
            else
            {
                const int64_t tb = ctx->h1_task_ptr[0][l], te = ctx->h1_task_ptr[0][l + 1];
                launch_chol_fwd_big( F, tasks + tb, te - tb, ctx->h1_level_wmax[l], ctx->stream );
            }
        }
        for( size_t l = nl; l-- > 0; )
        {
            if( ctx->h1_level_small_bwd[l] )
            {
                const int64_t b = ctx->h1_level_ptr[l], e = ctx->h1_level_ptr[l + 1];
                launch_chol_bwd_small( F, ctx->h1_sdesc.get() + b, e - b, ctx->h1_level_g[l], ctx->h1_level_rmax[l], ctx->stream );
            }
            else
            {
                const int64_t tb = ctx->h1_task_ptr[1][l], te = ctx->h1_task_ptr[1][l + 1];
                launch_chol_bwd_big( F, tasks + tb, te - tb, ctx->h1_level_rmax[l], ctx->h1_level_wmax[l] > CHOL_NARROW_WIDTH, ctx->stream );
            }
        }
        if( Xd ) { launch_chol_scatter( ctx->h1_n, ctx->h1_perm.get(), ctx->h1_xp.get(), beta, Xd, ldx, ctx->stream ); }
    }

    // Y = alpha P0 X + beta Y with P0 = H^-1 L_{2-s} H^-1.
    // Reference: TangentPointMetric0::multiply_preconditioner (src/Metrics/TangentPointMetric0.hpp:88-126).
    // Between the two H^1 solves the vectors stay in the factor's numbering: the first solve leaves its
    // solution in h1_xp, the pseudo-Laplacian's pre-operator reads it through pre_cols mapped by the
    // inverse permutation and its vertex gather writes row j = vertex perm[j] straight into h1_bp, the
    // second solve's right-hand side; the same values and sums as permuting back and forth in between.
    static bool pc0_fused( rp_ctx * ctx ) { return ctx->h1_n == ctx->V; }

    // The pseudo-Laplacian step of P0 in the factor's numbering: h1_bp = L h1_xp.
    static void pc0_middle( rp_ctx * ctx )
    {
        if( ctx->pre_cols_h1_key[0] != ctx->tree_version || ctx->pre_cols_h1_key[1] != ctx->h1_version )
        {
            const int64_t ne = ctx->T * ctx->size;
            ctx->pre_cols_h1.require( (size_t)std::max<int64_t>( 1, ne ) );
            launch_map_index( ne, ctx->h1_iperm.get(), ctx->pre_cols.get(), ctx->pre_cols_h1.get(), ctx->stream );
            RP_HIP( hipGetLastError() );
            ctx->pre_cols_h1_key[0] = ctx->tree_version;
            ctx->pre_cols_h1_key[1] = ctx->h1_version;
        }
        pl_apply( ctx, 1.0, ctx->h1_xp.get(), 3, 0.0, ctx->h1_bp.get(), 3, true, true, ctx->pre_cols_h1.get(), ctx->h1_perm.get() );
    }

    static void pc0_apply( rp_ctx * ctx, double alpha, const double * Xd, int64_t ldx, double beta, double * Yd, int64_t ldy )
    {
        const int64_t V = ctx->V;
        if( alpha == 0.0 || !pc0_fused( ctx ) )
        {
            ctx->pc_tmp.require( (size_t)(3 * V) );
            double * Z = ctx->pc_tmp.get();
            h1_solve( ctx, alpha, Xd, ldx, 0.0, Z, 3 );
            if( alpha != 0.0 ) { pl_apply( ctx, 1.0, Z, 3, 0.0, Z, 3, true, true ); }
            h1_solve( ctx, 1.0, Z, 3, beta, Yd, ldy );
            return;
        }
        h1_solve( ctx, alpha, Xd, ldx, 0.0, nullptr, 0 );
        pc0_middle( ctx );
        h1_solve( ctx, 1.0, nullptr, 0, beta, Yd, ldy );
    }

    // Exact diagonal of the assembled metric (cached per assembly).
    static void ensure_metric_diag( rp_ctx * ctx )
    {
        if( ctx->diag_version == ctx->metric_version ) { return; }
        ctx->dA.require( (size_t)ctx->V );
        ctx->dA_path.require( (size_t)(2 * ctx->C) );
        DiagArgs A{};
        A.V = ctx->V; A.C = ctx->C; A.size = ctx->size;
        A.have_near = ctx->nf_cols > 0; A.use_far = ctx->far_nnz > 0;
        A.vf_diag = ctx->have_vf_metric ? ctx->vf_diag.get() : nullptr;
        A.vp_ptr = ctx->vp_ptr.get(); A.vp_entry = ctx->vp_entry.get(); A.pre_vals = ctx->pre_vals.get();
        A.leaf_of_prim = ctx->leaf_of_prim.get(); A.leaf_begin = ctx->leaf_begin.get(); A.leaf_end = ctx->leaf_end.get();
        A.col_ptr = ctx->col_ptr.get(); A.col_prim = ctx->col_prim.get(); A.col_mask = ctx->col_mask.get(); A.val_ptr = ctx->val_ptr.get();
        A.nf_diag = ctx->nf_diag.get(); A.leaf_cluster = ctx->leaf_cluster.get(); A.C_parent = ctx->C_parent.get();
        A.C_begin = ctx->C_begin.get(); A.C_end = ctx->C_end.get(); A.ff_diag = ctx->ff_diag.get();
        launch_metric_diag( ctx->metric_precision == RP_MIXED, A, ctx->nf_vals.get(), ctx->dA_path.get(), ctx->dA.get(), ctx->stream );
        RP_HIP( hipGetLastError() );
        ctx->diag_version = ctx->metric_version;
    }

    // Weight w of the diagonal term. A negative setting estimates it per state as lambda_max(P0 A) / 4
    // by power iteration from the (smooth) vertex coordinates: the top of the spectrum of P0 A belongs to
    // smooth modes, while the rough modes that P0 under-weights get w diag(A)^-1 A ~ w.
    static constexpr int PC_POWER_STEPS = 12;

    // Vertex scalings sqrt(d_v) of the learned preconditioner from the node MLP (FP32 or BF16 weights).
    static void learned_scales( rp_ctx * ctx, bool bf16 )
    {
        const int F = ctx->lp_F;
        ctx->lp_lam.require( (size_t)ctx->C ); ctx->lp_leaf.require( (size_t)ctx->L ); ctx->lp_sqrtd.require( (size_t)ctx->V );
        ctx->lp_logd.require( (size_t)ctx->V ); ctx->lp_gp.require( (size_t)ctx->T );
        ctx->lp_params16.require( (size_t)mlp_param_count( F ) );
        launch_mlp_infer( bf16, ctx->C, F, ctx->lp_feats.get(), ctx->lp_params.get(), ctx->lp_params16.get(), ctx->lp_lam.get(), ctx->stream );
        launch_leaf_path_sum1( ctx->L, ctx->C, ctx->leaf_anc.get(), ctx->leaf_anc_len.get(), ctx->lp_lam.get(), ctx->lp_leaf.get(), ctx->stream );
        launch_vertex_scale( ctx->V, ctx->T, ctx->size, ctx->pre_cols.get(), ctx->vp_ptr.get(), ctx->vp_entry.get(), ctx->leaf_of_prim.get(),
                             ctx->lp_leaf.get(), ctx->lp_logd.get(), ctx->lp_gp.get(), ctx->lp_sqrtd.get(), ctx->stream );
        RP_HIP( hipGetLastError() );
    }

    // lambda_max(P0 A) / 4, cached per state of metric, pseudo-Laplacian and H1 factor.
    static double estimated_weight( rp_ctx * ctx )
    {
        const int64_t key[3] = { ctx->metric_version, ctx->pl_version, ctx->h1_version };
        if( key[0] == ctx->pc_weight_key[0] && key[1] == ctx->pc_weight_key[1] && key[2] == ctx->pc_weight_key[2] ) { return ctx->pc_weight_est; }
        const int64_t V = ctx->V;
        const size_t n3 = (size_t)(3 * V);
        ctx->cg_p.require( n3 ); ctx->cg_Ap.require( n3 ); ctx->cg_z.require( n3 );
        ctx->cg_scal.require( 32 ); ctx->cg_part.require( 3 * 1024 );
        double * v = ctx->cg_p.get(); double * t = ctx->cg_Ap.get(); double * u = ctx->cg_z.get();
        double * sc = ctx->cg_scal.get() + 24; double * part = ctx->cg_part.get();
        hipStream_t s = ctx->stream;
        launch_copy3( V, ctx->coords.get(), 3, v, 3, s );
        double lam = 0.0;
        for( int k = 0; k < PC_POWER_STEPS; ++k )
        {
            double nv[3], nu[3];
            launch_dot3( V, v, v, part, sc, s );
            RP_HIP( hipMemcpyAsync( nv, sc, sizeof(nv), hipMemcpyDeviceToHost, s ) );
            metric_apply( ctx, 1.0, v, 3, 0.0, t, 3, true, true, false );
            pc0_apply( ctx, 1.0, t, 3, 0.0, u, 3 );
            launch_dot3( V, u, u, part, sc, s );
            RP_HIP( hipMemcpyAsync( nu, sc, sizeof(nu), hipMemcpyDeviceToHost, s ) );
            RP_HIP( hipStreamSynchronize( s ) );
            lam = 0.0;
            for( int c = 0; c < 3; ++c ) { if( nv[c] > 0 ) { lam = std::max( lam, std::sqrt( nu[c] / nv[c] ) ); } }
            std::swap( v, u );
            launch_dot3( V, v, v, part, sc, s );
            launch_normalize3( V, v, sc, v, s );
        }
        ctx->pc_weight_est = 0.25 * lam;
        for( int k = 0; k < 3; ++k ) { ctx->pc_weight_key[k] = key[k]; }
        return ctx->pc_weight_est;
    }

    // Weight in use: the fixed setting or the estimate; a learned model multiplies its base weight (the
    // setting if positive, else the estimate) by exp(learned log w).
    static void update_weight( rp_ctx * ctx )
    {
        // setting > 0: fixed weight; setting < 0: |setting| times the estimate; 0: P0 alone (a learned model
        // then starts from the estimate)
        const double set = ctx->pc_weight_setting;
        double w;
        if( ctx->lp_enabled ) { w = ( set > 0.0 ? set : ( set < 0.0 ? -set : 1.0 ) * estimated_weight( ctx ) ) * std::exp( ctx->lp_log_w ); }
        else { w = set >= 0.0 ? set : -set * estimated_weight( ctx ); }
        ctx->pc_weight = w;
        if( w > 0.0 ) { ensure_metric_diag( ctx ); }
    }

    static void prepare_pc( rp_ctx * ctx )
    {
        if( ctx->lp_enabled && !ctx->lp_scale_ready ) { learned_scales( ctx, ctx->lp_bf16 ); ctx->lp_scale_ready = true; }
        update_weight( ctx );
    }

    // Y = alpha (P0 + w diag(A)^-1) X + beta Y (w = 0: the reference preconditioner). X and Y must not
    // overlap. Call prepare_pc first.
    static void pc2_apply( rp_ctx * ctx, double alpha, const double * Xd, int64_t ldx, double beta, double * Yd, int64_t ldy )
    {
        pc0_apply( ctx, alpha, Xd, ldx, beta, Yd, ldy );
        if( ctx->pc_weight > 0.0 && alpha != 0.0 ) { launch_jacobi_add( ctx->V, alpha * ctx->pc_weight, Xd, ldx, ctx->dA.get(), Yd, ldy, ctx->stream ); }
    }

    // Learned form: Y = alpha D^{1/2} (P0 + w diag(A)^-1) D^{1/2} X + beta Y.
    static void pc_apply( rp_ctx * ctx, double alpha, const double * Xd, int64_t ldx, double beta, double * Yd, int64_t ldy )
    {
        if( !ctx->lp_enabled ) { pc2_apply( ctx, alpha, Xd, ldx, beta, Yd, ldy ); return; }
        const int64_t V = ctx->V;
        if( alpha != 0.0 && pc0_fused( ctx ) )
        {
            // D^{1/2} folded into the first solve's gather, the diagonal term and the outer D^{1/2} into the
            // last solve's scatter (the same products as the separate steps below)
            launch_chol_gather_scaled( ctx->h1_n, ctx->h1_perm.get(), ctx->lp_sqrtd.get(), Xd, ldx, ctx->h1_bp.get(), ctx->stream );
            h1_solve( ctx, 1.0, nullptr, 0, 0.0, nullptr, 0 );
            pc0_middle( ctx );
            h1_solve( ctx, 1.0, nullptr, 0, 0.0, nullptr, 0 );
            const double cw = ctx->pc_weight > 0.0 ? 1.0 * ctx->pc_weight : 0.0;
            launch_chol_scatter_scaled( ctx->h1_n, ctx->h1_perm.get(), ctx->h1_xp.get(), ctx->lp_sqrtd.get(), Xd, ldx,
                                        cw, ctx->dA.get(), alpha, beta, Yd, ldy, ctx->stream );
            return;
        }
        ctx->lp_t1.require( (size_t)(3 * V) ); ctx->lp_t2.require( (size_t)(3 * V) );
        launch_row_scale( V, ctx->lp_sqrtd.get(), Xd, ldx, 1.0, 0.0, ctx->lp_t1.get(), 3, ctx->stream );
        pc2_apply( ctx, 1.0, ctx->lp_t1.get(), 3, 0.0, ctx->lp_t2.get(), 3 );
        launch_row_scale( V, ctx->lp_sqrtd.get(), ctx->lp_t2.get(), 3, alpha, beta, Yd, ldy, ctx->stream );
    }

    // Host staging of a V x 3 input/output pair for calls with host pointers.
    struct Staged
    {
        rp_ctx * ctx;
        bool dev;
        const double * X; int64_t ldx;
        double * Y; int64_t ldy;
        double * Yh; int64_t ldYh;
        std::vector<double> hx, hy;

        Staged( rp_ctx * c, bool device, const double * X_, int64_t ldX, double beta, double * Y_, int64_t ldY )
        : ctx(c), dev(device), X(X_), ldx(ldX), Y(Y_), ldy(ldY), Yh(Y_), ldYh(ldY)
        {
            if( dev ) { return; }
            const int64_t V = ctx->V;
            pack_rows( X_, ldX, V, hx );
            ctx->Xd.upload( hx.data(), hx.size(), ctx->stream );
            X = ctx->Xd.get(); ldx = 3;
            ctx->Yd.require( (size_t)(3 * V) );
            if( beta != 0.0 ) { pack_rows( Y_, ldY, V, hy ); ctx->Yd.upload( hy.data(), hy.size(), ctx->stream ); }
            Y = ctx->Yd.get(); ldy = 3;
        }

        // Downloads the result (host pointers) after the stream has finished.
        void finish()
        {
            if( dev ) { RP_HIP( hipStreamSynchronize( ctx->stream ) ); return; }
            hy.resize( (size_t)(3 * ctx->V) );
            ctx->Yd.download( hy.data(), hy.size(), ctx->stream );
            RP_HIP( hipStreamSynchronize( ctx->stream ) );
            unpack_rows( hy, ctx->V, Yh, ldYh );
        }
    };

    static void require_nrhs( int nrhs )
    {
        require( nrhs == 3, "nrhs = " + std::to_string(nrhs) + " != 3 = AMB_DIM. The current implementation only accepts nrhs = AMB_DIM. Doing nothing." );
    }

    // Preconditioned conjugate gradients on the three columns (independent scalars per column, zero
    // start). Stopping measure: |r|_2 / |b|_2, or with pc_norm the preconditioner norm
    // sqrt(<r, P r>) / sqrt(<b, P b>). b, x: device V x 3 (ld 3). Returns the iteration count.
    // Metric assembled at RP_MIXED_REFINED: the products A p use the FP32 copies of the near-field
    // values until the recursive residual passes; then r = b - A x with the FP64 values replaces it
    // (residual replacement, as in mixed-precision iterative refinement). If a column is still above
    // the tolerance, CG restarts from x (p = P r) and continues on the FP64 values.
    static int pcg( rp_ctx * ctx, const double * b, double * x, int max_iter, double tol, bool use_pc, bool pc_norm,
                    std::vector<double> & rel, bool & done, bool record )
    {
        const int64_t V = ctx->V;
        auto keep = [&]( const double * rv, const double * xv )
        {
            if( !record || ctx->hist_n >= ctx->hist_cap ) { return; }
            launch_copy3( V, rv, 3, ctx->hist_r.get() + (size_t)(3 * V) * ctx->hist_n, 3, ctx->stream );
            launch_copy3( V, xv, 3, ctx->hist_x.get() + (size_t)(3 * V) * ctx->hist_n, 3, ctx->stream );
            ++ctx->hist_n;
        };
        const size_t n3 = (size_t)(3 * V);
        hipStream_t s = ctx->stream;
        ctx->cg_r.require( n3 ); ctx->cg_p.require( n3 ); ctx->cg_Ap.require( n3 ); ctx->cg_z.require( n3 );
        ctx->cg_scal.require( 32 ); ctx->cg_part.require( 3 * 1024 );
        double * r = ctx->cg_r.get(); double * p = ctx->cg_p.get();
        double * Ap = ctx->cg_Ap.get(); double * z = ctx->cg_z.get();
        double * sc = ctx->cg_scal.get(); double * part = ctx->cg_part.get();
        double * rho[2] = { sc + 0, sc + 3 };
        double * pAp = sc + 6; double * rr = sc + 9; double * bb = sc + 12; double * active = sc + 15;
        double * rho0 = sc + 18;
        const double tol2 = tol * tol;
        auto precondition = [&]( const double * in, double * out )
        {
            if( use_pc ) { pc_apply( ctx, 1.0, in, 3, 0.0, out, 3 ); }
            else         { launch_copy3( V, in, 3, out, 3, s ); }
        };
        // x = 0, r = b, z = P r, p = z, rho = <r, z>
        launch_fill( x, (int64_t)n3, 0.0, s );
        launch_copy3( V, b, 3, r, 3, s );
        keep( r, x );
        launch_dot3( V, r, r, part, bb, s );
        precondition( r, z );
        launch_copy3( V, z, 3, p, 3, s );
        launch_dot3( V, r, z, part, rho[0], s );
        launch_copy3( 1, rho[0], 3, rho0, 3, s );
        const double * ref = pc_norm ? rho0 : bb;
        launch_cg_check( ref, ref, 0.0, active, s );
        double h_m[3], h_ref[3];
        RP_HIP( hipMemcpyAsync( h_ref, ref, sizeof(h_ref), hipMemcpyDeviceToHost, s ) );
        RP_HIP( hipStreamSynchronize( s ) );
        rel.assign( 3, 0.0 );
        auto all_done = [&]( const double * m )
        {
            bool ok = true;
            for( int k = 0; k < 3; ++k )
            {
                rel[(size_t)k] = h_ref[k] > 0 ? std::sqrt( std::max( 0.0, m[k] ) / h_ref[k] ) : 0.0;
                if( !( m[k] <= tol2 * h_ref[k] ) ) { ok = false; }   // a NaN residual is not converged
            }
            return ok;
        };
        // the guard: at iteration guard_iter, stop when the largest relative residual is above guard_ratio
        ctx->guard_tripped = false;
        auto guard_hit = [&]( const int k ) -> bool
        {
            if( ctx->guard_iter <= 0 || k != ctx->guard_iter ) { return false; }
            bool over = false;
            for( int c = 0; c < 3; ++c ) { if( !( rel[(size_t)c] <= ctx->guard_ratio ) ) { over = true; } }   // NaN trips it
            ctx->guard_tripped = over;
            return ctx->guard_tripped;
        };
        int it = 0;
        done = true;   // nothing to do when every column of B vanishes
        for( int k = 0; k < 3; ++k ) { if( h_ref[k] > 0 ) { done = false; } }
        int cur = 0;
        bool inner32 = ctx->metric_precision == RP_MIXED_REFINED;
        // r = b - A x on the FP64 values; a new measure per column; false if a column is still above
        // the tolerance (then z = P r and rho[cur] = <r, z> are set for the restart)
        auto replace_residual = [&]() -> bool
        {
            launch_copy3( V, b, 3, r, 3, s );
            metric_apply( ctx, -1.0, x, 3, 1.0, r, 3, true, true, false );
            if( !pc_norm )
            {
                launch_dot3( V, r, r, part, rr, s );
                RP_HIP( hipMemcpyAsync( h_m, rr, sizeof(h_m), hipMemcpyDeviceToHost, s ) );
                launch_cg_check( rr, bb, tol2, active, s );
                RP_HIP( hipStreamSynchronize( s ) );
                if( all_done( h_m ) ) { return true; }
                precondition( r, z );
                launch_dot3( V, r, z, part, rho[cur], s );
            }
            else
            {
                precondition( r, z );
                launch_dot3( V, r, z, part, rho[cur], s );
                RP_HIP( hipMemcpyAsync( h_m, rho[cur], sizeof(h_m), hipMemcpyDeviceToHost, s ) );
                launch_cg_check( rho[cur], rho0, tol2, active, s );
                RP_HIP( hipStreamSynchronize( s ) );
                if( all_done( h_m ) ) { return true; }
            }
            launch_copy3( V, z, 3, p, 3, s );
            return false;
        };
        // While the host waits for an iteration's residual, the stream already runs the next
        // preconditioner application when the residual history says this iteration will not converge
        // (the last relative residual times the last reduction ratio still above twice the tolerance):
        // the device work and its order are the same either way, only the host's wait moves, so the
        // GPU does not drain and refill once per iteration. A wrong guess costs one unused
        // preconditioner application.
        double rel_prev = 1.0, ratio_prev = 1.0;
        double * h_pin = ctx->h_pinned;
        while( !done && it < max_iter )
        {
            metric_apply( ctx, 1.0, p, 3, 0.0, Ap, 3, true, true, false, true, inner32 );
            launch_dot3( V, p, Ap, part, pAp, s );
            launch_cg_update( V, x, r, p, Ap, rho[cur], pAp, active, part, rr, s );
            keep( r, x );
            if( !pc_norm )
            {
                RP_HIP( hipMemcpyAsync( h_pin, rr, 3 * sizeof(double), hipMemcpyDeviceToHost, s ) );
                launch_cg_check( rr, bb, tol2, active, s );
                RP_HIP( hipEventRecord( ctx->ev_check, s ) );
                ++it;
                const bool ahead = it < max_iter && rel_prev * std::min( 1.0, ratio_prev ) > 2.0 * tol;
                if( ahead )
                {
                    precondition( r, z );
                    launch_dot3( V, r, z, part, rho[cur ^ 1], s );
                }
                RP_HIP( hipEventSynchronize( ctx->ev_check ) );
                for( int k = 0; k < 3; ++k ) { h_m[k] = h_pin[k]; }
                done = all_done( h_m );
                {
                    double worst = 0.0;
                    for( int k = 0; k < 3; ++k ) { if( !( rel[(size_t)k] <= tol ) ) { worst = std::max( worst, std::isfinite( rel[(size_t)k] ) ? rel[(size_t)k] : 1.0 ); } }
                    ratio_prev = rel_prev > 0.0 ? worst / rel_prev : 1.0;
                    rel_prev = worst;
                }
                if( !done && guard_hit( it ) ) { break; }
                if( done && inner32 )
                {
                    inner32 = false;
                    done = replace_residual();
                    if( done || it >= max_iter ) { break; }
                    continue;
                }
                if( done || it >= max_iter ) { break; }
                if( !ahead )
                {
                    precondition( r, z );
                    launch_dot3( V, r, z, part, rho[cur ^ 1], s );
                }
            }
            else
            {
                precondition( r, z );
                launch_dot3( V, r, z, part, rho[cur ^ 1], s );
                RP_HIP( hipMemcpyAsync( h_m, rho[cur ^ 1], sizeof(h_m), hipMemcpyDeviceToHost, s ) );
                launch_cg_check( rho[cur ^ 1], rho0, tol2, active, s );
                RP_HIP( hipStreamSynchronize( s ) );
                ++it;
                done = all_done( h_m );
                if( !done && guard_hit( it ) ) { break; }
                if( done && inner32 )
                {
                    inner32 = false;
                    done = replace_residual();
                    if( done || it >= max_iter ) { break; }
                    continue;
                }
                if( done || it >= max_iter ) { break; }
            }
            launch_cg_dir( V, p, z, rho[cur ^ 1], rho[cur], active, s );
            cur ^= 1;
        }
        if( inner32 && !done )
        {
            // iteration cap reached on the FP32 copies: judge and report the FP64 residual
            done = replace_residual();
        }
        return it;
    }

    constexpr int GMRES_RESTART = 100;

    // Restarted GMRES(m), left preconditioned (Saad & Schultz 1986): minimises |P (b - A x)|_2 over
    // x in x0 + K_m(P A, P r0) for each of the three columns, Arnoldi with classical Gram-Schmidt
    // applied twice (each pass is one batched inner-product kernel and one batched update), Givens
    // rotations on the host. b: device V x 3 (ld 3), x: device V x 3 (ld 3, overwritten, zero start).
    // Returns the number of Arnoldi steps; rel receives |P r|_2 / |P b|_2 per column.
    // Reference: MetricDimRestricted::Solve with GMRES<3,Real,Int,Side::Left>
    // (src/Metrics/MetricDimRestricted.hpp:232-289).
    static int gmres( rp_ctx * ctx, const double * b, double * x, int max_iter, double tol, bool use_pc, std::vector<double> & rel, bool & converged )
    {
        const int64_t V = ctx->V;
        const int64_t n3 = 3 * V;
        const int m = std::max( 1, std::min( max_iter, GMRES_RESTART ) );
        ctx->gm_V.require( (size_t)( (m + 1) * n3 ) );
        ctx->gm_w.require( (size_t)n3 );
        ctx->gm_t.require( (size_t)n3 );
        ctx->gm_small.require( (size_t)( 4 * 3 * (m + 2) ) );
        ctx->gm_part.require( (size_t)( (m + 1) * MDOT_PART ) );
        double * Vb = ctx->gm_V.get(); double * w = ctx->gm_w.get(); double * t = ctx->gm_t.get();
        double * h1 = ctx->gm_small.get(); double * h2 = h1 + 3 * (m + 2); double * nrm = h2 + 3 * (m + 2); double * y = nrm + 3 * (m + 2);
        double * part = ctx->gm_part.get();
        hipStream_t s = ctx->stream;
        auto precondition = [&]( const double * in, double * out )
        {
            if( use_pc ) { pc_apply( ctx, 1.0, in, 3, 0.0, out, 3 ); }
            else         { launch_copy3( V, in, 3, out, 3, s ); }
        };
        launch_fill( x, n3, 0.0, s );
        precondition( b, w );
        launch_dot3( V, w, w, part, nrm, s );
        double hn[3];
        RP_HIP( hipMemcpyAsync( hn, nrm, sizeof(hn), hipMemcpyDeviceToHost, s ) );
        RP_HIP( hipStreamSynchronize( s ) );
        double beta0[3];
        for( int k = 0; k < 3; ++k ) { beta0[k] = std::sqrt( hn[k] ); }
        rel.assign( 3, 0.0 );
        converged = ( beta0[0] == 0 && beta0[1] == 0 && beta0[2] == 0 );
        int it = 0;
        std::vector<double> R( (size_t)(3 * m * (m + 1)) ), g( (size_t)(3 * (m + 1)) ), cs( (size_t)(3 * m) ), sn( (size_t)(3 * m) );
        std::vector<double> hh( (size_t)(3 * (m + 2)) ), hh2( (size_t)(3 * (m + 2)) ), yy( (size_t)(3 * (m + 1)) );
        double beta[3] = { beta0[0], beta0[1], beta0[2] };
        while( !converged && it < max_iter )
        {
            launch_normalize3( V, w, nrm, Vb, s );
            std::fill( g.begin(), g.end(), 0.0 );
            for( int k = 0; k < 3; ++k ) { g[(size_t)k] = beta[k]; }
            int j = 0;
            for( ; j < m && it < max_iter; ++j )
            {
                double * vj = Vb + n3 * j;
                double * vn = Vb + n3 * (j + 1);
                metric_apply( ctx, 1.0, vj, 3, 0.0, t, 3, true, true, false );
                precondition( t, w );
                launch_mdot( V, Vb, n3, j + 1, w, part, h1, 0, s );
                launch_maxpy( V, Vb, n3, j + 1, h1, -1.0, w, s );
                launch_mdot( V, Vb, n3, j + 1, w, part, h2, 0, s );
                launch_maxpy( V, Vb, n3, j + 1, h2, -1.0, w, s );
                launch_dot3( V, w, w, part, nrm, s );
                launch_normalize3( V, w, nrm, vn, s );
                RP_HIP( hipMemcpyAsync( hh.data(), h1, sizeof(double) * 3 * (j + 1), hipMemcpyDeviceToHost, s ) );
                RP_HIP( hipMemcpyAsync( hh2.data(), h2, sizeof(double) * 3 * (j + 1), hipMemcpyDeviceToHost, s ) );
                RP_HIP( hipMemcpyAsync( hn, nrm, sizeof(hn), hipMemcpyDeviceToHost, s ) );
                RP_HIP( hipStreamSynchronize( s ) );
                ++it;
                bool all = true;
                for( int k = 0; k < 3; ++k )
                {
                    double * Hc = &R[(size_t)(k * m * (m + 1) + j * (m + 1))];   // column j of column k's Hessenberg
                    for( int i = 0; i <= j; ++i ) { Hc[i] = hh[(size_t)(3 * i + k)] + hh2[(size_t)(3 * i + k)]; }
                    Hc[j + 1] = std::sqrt( hn[k] );
                    for( int i = 0; i < j; ++i )
                    {
                        const double c_ = cs[(size_t)(k * m + i)], s_ = sn[(size_t)(k * m + i)];
                        const double a = Hc[i], bq = Hc[i + 1];
                        Hc[i] = c_ * a + s_ * bq;
                        Hc[i + 1] = -s_ * a + c_ * bq;
                    }
                    const double d = std::hypot( Hc[j], Hc[j + 1] );
                    const double c_ = d > 0 ? Hc[j] / d : 1.0, s_ = d > 0 ? Hc[j + 1] / d : 0.0;
                    cs[(size_t)(k * m + j)] = c_; sn[(size_t)(k * m + j)] = s_;
                    Hc[j] = d; Hc[j + 1] = 0.0;
                    const double gj = g[(size_t)(3 * j + k)];   // g[3 i + k]: rotated right-hand side
                    g[(size_t)(3 * (j + 1) + k)] = -s_ * gj;
                    g[(size_t)(3 * j + k)] = c_ * gj;
                    rel[(size_t)k] = beta0[k] > 0 ? std::abs( g[(size_t)(3 * (j + 1) + k)] ) / beta0[k] : 0.0;
                    if( !( rel[(size_t)k] <= tol ) ) { all = false; }   // a NaN residual is not converged
                }
                if( all ) { converged = true; ++j; break; }
            }
            // x += V y with R y = g (upper triangular, per column)
            const int nj = j;
            for( int k = 0; k < 3; ++k )
            {
                for( int i = nj - 1; i >= 0; --i )
                {
                    double acc = g[(size_t)(3 * i + k)];
                    for( int l = i + 1; l < nj; ++l ) { acc -= R[(size_t)(k * m * (m + 1) + l * (m + 1) + i)] * yy[(size_t)(3 * l + k)]; }
                    const double rii = R[(size_t)(k * m * (m + 1) + i * (m + 1) + i)];
                    yy[(size_t)(3 * i + k)] = rii != 0.0 ? acc / rii : 0.0;
                }
            }
            RP_HIP( hipMemcpyAsync( y, yy.data(), sizeof(double) * 3 * nj, hipMemcpyHostToDevice, s ) );
            launch_maxpy( V, Vb, n3, nj, y, 1.0, x, s );
            if( converged || it >= max_iter ) { break; }
            // restart: w = P (b - A x)
            metric_apply( ctx, 1.0, x, 3, 0.0, t, 3, true, true, false );
            launch_axpby3( V, 1.0, b, 3, -1.0, t, 3, s );
            precondition( t, w );
            launch_dot3( V, w, w, part, nrm, s );
            RP_HIP( hipMemcpyAsync( hn, nrm, sizeof(hn), hipMemcpyDeviceToHost, s ) );
            RP_HIP( hipStreamSynchronize( s ) );
            for( int k = 0; k < 3; ++k ) { beta[k] = std::sqrt( hn[k] ); }
        }
        RP_HIP( hipStreamSynchronize( s ) );
        return it;
    }
}

using namespace rp;

extern "C"
{

int rp_pl_compute( rp_ctx * ctx, double s, int precision )
{
    return guarded( "rp_pl_compute", [&]()
    {
        require( ctx != nullptr && ctx->have_data && ctx->have_near && ctx->have_far, "mesh, tree, patterns and data must be uploaded", RP_ENOTREADY );
        require( precision == RP_FP64 || precision == RP_MIXED || precision == RP_PL_BF16, "unknown precision" );
        require( std::isfinite( s ), "s must be finite" );
        RP_HIP( hipSetDevice( ctx->device ) );
        const int n = ctx->dom_dim;
        const int64_t T = ctx->T, C = ctx->C, L = ctx->L;
        const double s_exp = -0.5 * n - s;
        const Expo ex = make_expo( 0.0, 0.0, n, s_exp );
        ctx->pl_nf_vals.require( (size_t)std::max<int64_t>( 1, ctx->nf_slots ) );
        ctx->pl_nf_diag.require( (size_t)T );
        ctx->pl_ff_vals.require( (size_t)std::max<int64_t>( 1, ctx->far_nnz ) );
        ctx->pl_ff_diag.require( (size_t)C );
        const bool have_nf = ctx->nf_cols > 0, have_ff = ctx->far_nnz > 0;
        EventScope ev( ctx, 9 );
        if( !have_nf ) { launch_fill( ctx->pl_nf_diag.get(), T, 0.0, ctx->stream ); }
        if( !have_ff ) { launch_fill( ctx->pl_ff_diag.get(), C, 0.0, ctx->stream ); }
        NFArgs A{};
        A.L = L; A.T = T;
        A.leaf_begin = ctx->leaf_begin.get(); A.leaf_end = ctx->leaf_end.get();
        A.col_ptr = ctx->col_ptr.get(); A.col_prim = ctx->col_prim.get(); A.col_mask = ctx->col_mask.get(); A.val_ptr = ctx->val_ptr.get();
        A.a = ctx->p_a.get(); A.x = ctx->p_x.get(); A.rec = ctx->p_rec.get(); A.ex = ex;
        A.vals = ctx->pl_nf_vals.get(); A.diag = ctx->pl_nf_diag.get();
        FFArgs B{};
        B.C = C; B.rp = ctx->far_rp.get(); B.ci = ctx->far_ci.get();
        B.a = ctx->c_a.get(); B.x = ctx->c_x.get(); B.ex = ex;
        B.vals = ctx->pl_ff_vals.get(); B.diag = ctx->pl_ff_diag.get();
        launch_pl_assemble( precision, A, B, have_nf, have_ff, ctx->stream );
        ctx->have_pl_vf = false;
        double st[3] = { 0, 0, 0 };
        if( ctx->vf_nnz > 0 )
        {
            // very-near field in FP64 (PseudoLaplacian_Traversor::VF_Compute, max_refinement 20 as in
            // PseudoLaplacian_Traversor.hpp:187)
            const int64_t nnz = ctx->vf_nnz;
            ctx->vf_rec.require( (size_t)(VF_REC * nnz) );
            ctx->pl_vf_vals.require( (size_t)nnz ); ctx->pl_vf_diag.require( (size_t)T );
            ctx->scalar.require( (size_t)(1 + REDUCE_BLOCKS + 3) );
            launch_fill( ctx->pl_vf_diag.get(), T, 0.0, ctx->stream );
            VFArgs V{};
            V.nnz = nnz; V.T = T; V.row = ctx->vf_row.get(); V.col = ctx->vf_ci.get();
            V.P_ordering = ctx->P_ordering.get(); V.simplices = ctx->simplices.get(); V.coords = ctx->coords.get();
            V.a = ctx->p_a.get(); V.P = ctx->p_P.get();
            V.theta2 = ctx->vf_theta * ctx->vf_theta; V.max_ref = 20;
            V.ex = ex; V.pl_sexp = s_exp; V.cap = VF_LANE_EVALUATION_LIMIT; V.rec = ctx->vf_rec.get();
            launch_vf_pairs( n, 1, V, ctx->stream );
            VFRowArgs W{};
            W.nrows = ctx->vf_nrows; W.T = T; W.rows = ctx->vf_rows.get(); W.rp = ctx->vf_rp.get(); W.dg = ctx->vf_dg.get();
            W.tr = ctx->vf_tr.get(); W.rec = ctx->vf_rec.get(); W.vals = ctx->pl_vf_vals.get(); W.diag = ctx->pl_vf_diag.get();
            launch_vf_rows( n, 1, W, ctx->stream );
            double * vf_stat = ctx->scalar.get() + 1 + REDUCE_BLOCKS;
            launch_vf_stats( nnz, ctx->vf_rec.get(), vf_stat, ctx->stream );
            RP_HIP( hipMemcpyAsync( st, vf_stat, sizeof(st), hipMemcpyDeviceToHost, ctx->stream ) );
        }
        ev.stop();
        RP_HIP( hipGetLastError() );
        RP_HIP( hipStreamSynchronize( ctx->stream ) );
        finish_timing( ctx, 9 );
        if( st[2] != 0.0 ) { ctx->have_pl = false; }
        require( st[2] == 0.0, "very-near-field subdivision exceeded 2^24 evaluations on one level-1 sub-pair "
                 "(primitives almost touching); refine the mesh", RP_EUNSUPPORTED );
        ctx->have_pl_vf = ctx->vf_nnz > 0;
        ctx->have_pl = true;
        ++ctx->pl_version;
        ctx->pl_precision = precision;
        ctx->pl_s = s;
    });
}

int rp_download_pl_vf( rp_ctx * ctx, double * vf_vals, double * vf_diag )
{
    return guarded( "rp_download_pl_vf", [&]()
    {
        require( ctx != nullptr && ctx->have_pl, "no pseudo-Laplacian computed", RP_ENOTREADY );
        RP_HIP( hipSetDevice( ctx->device ) );
        if( !ctx->have_pl_vf )
        {
            if( vf_diag ) { std::fill( vf_diag, vf_diag + ctx->T, 0.0 ); }
            return;
        }
        if( vf_vals ) { ctx->pl_vf_vals.download( vf_vals, (size_t)ctx->vf_nnz, ctx->stream ); }
        if( vf_diag ) { ctx->pl_vf_diag.download( vf_diag, (size_t)ctx->T, ctx->stream ); }
        RP_HIP( hipStreamSynchronize( ctx->stream ) );
    });
}

int rp_download_pl( rp_ctx * ctx, double * nf_vals, double * nf_diag, double * ff_vals, double * ff_diag )
{
    return guarded( "rp_download_pl", [&]()
    {
        require( ctx != nullptr && ctx->have_pl, "no pseudo-Laplacian computed", RP_ENOTREADY );
        RP_HIP( hipSetDevice( ctx->device ) );
        // bytes per stored value: FP64 8, FP32 4, BF16 2 (near field only; its far field is FP32)
        auto get_vals = [&]( const DBuf<double> & buf, int64_t cnt, double * out, int width )
        {
            if( !out || cnt == 0 ) { return; }
            if( width == 8 ) { buf.download( out, (size_t)cnt, ctx->stream ); RP_HIP( hipStreamSynchronize( ctx->stream ) ); return; }
            std::vector<uint8_t> raw( (size_t)cnt * (size_t)width );
            RP_HIP( hipMemcpyAsync( raw.data(), buf.get(), raw.size(), hipMemcpyDeviceToHost, ctx->stream ) );
            RP_HIP( hipStreamSynchronize( ctx->stream ) );
            for( size_t k = 0; k < (size_t)cnt; ++k )
            {
                float f = 0.0f;
                if( width == 4 ) { std::memcpy( &f, raw.data() + 4 * k, 4 ); }
                else
                {
                    uint16_t h = 0;
                    std::memcpy( &h, raw.data() + 2 * k, 2 );
                    const uint32_t bits = (uint32_t)h << 16;   // BF16 is the top half of an FP32
                    std::memcpy( &f, &bits, 4 );
                }
                out[k] = f;
            }
        };
        const int p = ctx->pl_precision;
        get_vals( ctx->pl_nf_vals, ctx->nf_cols > 0 ? ctx->nf_slots : 0, nf_vals, p == RP_FP64 ? 8 : p == RP_MIXED ? 4 : 2 );
        get_vals( ctx->pl_ff_vals, ctx->far_nnz, ff_vals, p == RP_FP64 ? 8 : 4 );
        if( nf_diag ) { ctx->pl_nf_diag.download( nf_diag, (size_t)ctx->T, ctx->stream ); }
        if( ff_diag ) { ctx->pl_ff_diag.download( ff_diag, (size_t)ctx->C, ctx->stream ); }
        RP_HIP( hipStreamSynchronize( ctx->stream ) );
    });
}

int rp_pl_multiply( rp_ctx * ctx, double alpha, const double * X, int64_t ldX, double beta, double * Y, int64_t ldY,
                    int nrhs, int nf_flag, int ff_flag, int ptr_flags )
{
    return guarded( "rp_pl_multiply", [&]()
    {
        require( ctx != nullptr, "ctx is NULL" );
        require_nrhs( nrhs );
        require( ctx->have_pl, "pseudo-Laplacian not assembled (call rp_pl_compute)", RP_ENOTREADY );
        require( X != nullptr && Y != nullptr && ldX >= 3 && ldY >= 3, "X and Y must be V x 3 with ld >= 3" );
        RP_HIP( hipSetDevice( ctx->device ) );
        Staged st( ctx, ptr_flags & RP_DEVICE_PTR, X, ldX, beta, Y, ldY );
        pl_apply( ctx, alpha, st.X, st.ldx, beta, st.Y, st.ldy, nf_flag != 0, ff_flag != 0 );
        RP_HIP( hipGetLastError() );
        st.finish();
    });
}

namespace
{
    // The factor's values on the structure of its tasks: the uploaded blocks hold L_off below inv(L_ss);
    // the device keeps M = L_off inv(L_ss) there, and row-major copies of the blocks of every level whose
    // backward runs lane groups (same offsets as in h1_vals; descriptors h1_sdesc, levels lptr).
    void h1_load_values( rp_ctx * ctx, const double * vals, int64_t nvals, int64_t mb, int64_t me,
                         const std::vector<int64_t> & lptr, const std::vector<char> & small_bwd, int64_t vt_extent )
    {
        rp::DBuf<double> lin;
        lin.upload( vals, (size_t)nvals, ctx->stream );
        ctx->h1_vals.require( (size_t)nvals );
        launch_chol_form_m( ctx->h1_tasks.get() + mb, me - mb, lin.get(), ctx->h1_vals.get(), ctx->stream );
        RP_HIP( hipGetLastError() );
        ctx->h1_vals_t.require( (size_t)vt_extent );
        const int64_t nl = (int64_t)small_bwd.size();
        for( int64_t l = 0; l < nl; ++l )
        {
            if( !small_bwd[(size_t)l] ) { continue; }
            launch_chol_transpose( ctx->h1_sdesc.get() + lptr[(size_t)l], lptr[(size_t)l + 1] - lptr[(size_t)l], ctx->h1_vals.get(), ctx->h1_vals_t.get(), ctx->stream );
        }
        RP_HIP( hipGetLastError() );
        RP_HIP( hipStreamSynchronize( ctx->stream ) );
    }
}

int rp_update_h1_values( rp_ctx * ctx, const double * vals )
{
    return guarded( "rp_update_h1_values", [&]()
    {
        require( ctx != nullptr, "ctx is NULL" );
        require( ctx->have_h1_structure, "no H1 factor structure uploaded (call rp_upload_h1_factor)", RP_ENOTREADY );
        require( vals != nullptr, "missing factor values" );
        RP_HIP( hipSetDevice( ctx->device ) );
        const size_t nl = ctx->h1_level_small_bwd.size();
        h1_load_values( ctx, vals, ctx->h1_nvals, ctx->h1_task_ptr[2][0], ctx->h1_task_ptr[2][nl], ctx->h1_level_ptr,
                        ctx->h1_level_small_bwd, ctx->h1_vt_extent );
        ctx->have_h1 = true;
        ++ctx->h1_version;
    });
}

int rp_upload_h1_factor( rp_ctx * ctx, int32_t n, int32_t ns, const int32_t * perm, const int32_t * sn_begin,
                         const int64_t * row_ptr, const int32_t * rows, const int64_t * val_ptr, const double * vals,
                         const int32_t * sn_parent )
{
    return guarded( "rp_upload_h1_factor", [&]()
    {
        require( ctx != nullptr && ctx->V > 0, "upload the mesh first", RP_ENOTREADY );
        require( n == ctx->V, "factor size differs from the vertex count" );
        require( ns > 0 && perm && sn_begin && row_ptr && rows && val_ptr && vals && sn_parent, "missing factor arrays" );
        RP_HIP( hipSetDevice( ctx->device ) );
        require( sn_begin[0] == 0 && sn_begin[ns] == n && row_ptr[0] == 0 && val_ptr[0] == 0, "bad supernode pointers" );
        {
            std::vector<char> seen( (size_t)n, 0 );
            for( int32_t j = 0; j < n; ++j )
            {
                require( perm[j] >= 0 && perm[j] < n && !seen[(size_t)perm[j]], "perm is not a permutation" );
                seen[(size_t)perm[j]] = 1;
            }
        }
        std::vector<int32_t> height( (size_t)ns, 0 );
        std::vector<int32_t> kid_ptr( (size_t)ns + 1, 0 );
        for( int32_t s = 0; s < ns; ++s )
        {
            const int32_t w = sn_begin[s + 1] - sn_begin[s];
            const int64_t r = row_ptr[s + 1] - row_ptr[s];
            require( w >= 1 && w <= CHOL_MAX_WIDTH, "supernode widths must lie in 1..2048", RP_EUNSUPPORTED );
            require( r >= w && r <= n, "bad supernode row count" );
            require( val_ptr[s + 1] - val_ptr[s] == r * w, "val_ptr must hold r x w blocks" );
            const int32_t * R = rows + row_ptr[s];
            for( int32_t a = 0; a < w; ++a ) { require( R[a] == sn_begin[s] + a, "row sets must start with the supernode's columns" ); }
            for( int64_t q = 1; q < r; ++q ) { require( R[q] > R[q - 1] && R[q] < n, "row sets must be sorted and in range" ); }
            const int32_t p = sn_parent[s];
            require( p == -1 || (p > s && p < ns), "supernodes must be in postorder (parent after child)" );
            if( p >= 0 ) { ++kid_ptr[(size_t)p + 1]; height[(size_t)p] = std::max( height[(size_t)p], height[(size_t)s] + 1 ); }
            else { require( r == w, "a root supernode cannot have update rows" ); }
        }
        for( int32_t s = 0; s < ns; ++s ) { kid_ptr[(size_t)s + 1] += kid_ptr[(size_t)s]; }
        std::vector<int32_t> kids( (size_t)std::max( 1, kid_ptr.back() ) );
        {
            std::vector<int32_t> pos( kid_ptr.begin(), kid_ptr.end() - 1 );
            for( int32_t s = 0; s < ns; ++s ) { if( sn_parent[s] >= 0 ) { kids[(size_t)pos[(size_t)sn_parent[s]]++] = s; } }
        }
        // cmap: position of each update row of a child in its parent's (sorted) row list
        const int64_t nrows = row_ptr[ns];
        std::vector<int32_t> cmap( (size_t)nrows, -1 );
        for( int32_t s = 0; s < ns; ++s )
        {
            const int32_t p = sn_parent[s];
            if( p < 0 ) { continue; }
            const int32_t w = sn_begin[s + 1] - sn_begin[s];
            const int32_t * Rp = rows + row_ptr[p];
            const int64_t rpn = row_ptr[p + 1] - row_ptr[p];
            for( int64_t q = row_ptr[s] + w; q < row_ptr[s + 1]; ++q )
            {
                const int32_t * it = std::lower_bound( Rp, Rp + rpn, rows[q] );
                require( it != Rp + rpn && *it == rows[q], "a child's update row is missing from its parent's row set" );
                cmap[(size_t)q] = (int32_t)( it - Rp );
            }
        }
        // pull lists: per front row, the child update rows mapping to it, children in increasing order
        require( nrows < (int64_t)INT32_MAX, "factor has too many row entries for 32-bit pull lists", RP_EUNSUPPORTED );
        std::vector<int32_t> pull_ptr( (size_t)nrows + 1, 0 );
        for( int32_t s = 0; s < ns; ++s )
        {
            const int32_t p = sn_parent[s];
            if( p < 0 ) { continue; }
            const int32_t w = sn_begin[s + 1] - sn_begin[s];
            for( int64_t q = row_ptr[s] + w; q < row_ptr[s + 1]; ++q ) { ++pull_ptr[(size_t)( row_ptr[p] + cmap[(size_t)q] ) + 1]; }
        }
        for( int64_t q = 0; q < nrows; ++q ) { pull_ptr[(size_t)q + 1] += pull_ptr[(size_t)q]; }
        std::vector<int32_t> pull_src( (size_t)std::max<int64_t>( 1, pull_ptr.back() ) );
        {
            std::vector<int32_t> pos( pull_ptr.begin(), pull_ptr.end() - 1 );
            for( int32_t p = 0; p < ns; ++p )
            {
                for( int32_t kk = kid_ptr[(size_t)p]; kk < kid_ptr[(size_t)p + 1]; ++kk )
                {
                    const int32_t c = kids[(size_t)kk];
                    const int32_t wc = sn_begin[c + 1] - sn_begin[c];
                    for( int64_t q = row_ptr[c] + wc; q < row_ptr[c + 1]; ++q )
                    {
                        const int64_t tgt = row_ptr[p] + cmap[(size_t)q];
                        pull_src[(size_t)pos[(size_t)tgt]++] = (int32_t)q;
                    }
                }
            }
        }
        // the sources of a front row in one 8-byte entry (almost every row has at most two), so that the
        // front merge waits for one index load instead of two
        std::vector<int32_t> pull2( (size_t)( 2 * nrows ) );
        for( int64_t q = 0; q < nrows; ++q )
        {
            const int32_t e0 = pull_ptr[(size_t)q], e1 = pull_ptr[(size_t)q + 1];
            int32_t * o = pull2.data() + 2 * q;
            if( e1 - e0 <= 2 )
            {
                o[0] = ( e1 > e0 ) ? pull_src[(size_t)e0] : -1;
                o[1] = ( e1 > e0 + 1 ) ? pull_src[(size_t)e0 + 1] : -1;
            }
            else { o[0] = -2 - e0; o[1] = e1; }
        }
        int32_t nl = 0;
        for( auto h : height ) { nl = std::max( nl, h + 1 ); }
        std::vector<int64_t> lptr( (size_t)nl + 1, 0 );
        for( auto h : height ) { ++lptr[(size_t)h + 1]; }
        for( int32_t l = 0; l < nl; ++l ) { lptr[(size_t)l + 1] += lptr[(size_t)l]; }
        std::vector<int32_t> lnodes( (size_t)ns );
        std::vector<int> wmax( (size_t)nl, 1 ), rmax( (size_t)nl, 1 ), gsz( (size_t)nl, 64 );
        std::vector<int64_t> rsum( (size_t)nl, 0 );
        {
            std::vector<int64_t> pos( lptr.begin(), lptr.end() - 1 );
            for( int32_t s = 0; s < ns; ++s )
            {
                const int32_t h = height[(size_t)s];
                lnodes[(size_t)pos[(size_t)h]++] = s;
                wmax[(size_t)h] = std::max( wmax[(size_t)h], sn_begin[s + 1] - sn_begin[s] );
                rmax[(size_t)h] = std::max( rmax[(size_t)h], (int)( row_ptr[s + 1] - row_ptr[s] ) );
                rsum[(size_t)h] += row_ptr[s + 1] - row_ptr[s];
            }
            // small levels: G lanes per supernode, the power of two covering the level's mean row count,
            // at least 8; rows (forward) and columns (backward) beyond G loop
            for( int32_t l = 0; l < nl; ++l )
            {
                const int64_t cnt = lptr[(size_t)l + 1] - lptr[(size_t)l];
                const int64_t need = std::max<int64_t>( (int64_t)8, cnt > 0 ? ( rsum[(size_t)l] + cnt - 1 ) / cnt : 1 );
                int g = 8;
                while( g < need && g < 64 ) { g <<= 1; }
                gsz[(size_t)l] = g;
            }
        }
        // Levels whose supernodes all have at most CHOL_SMALL_ROWS rows run a lane group per supernode
        // (the backward also levels of at most CHOL_SMALL_ROWS_MANY rows with at least
        // CHOL_MANY_SUPERNODES supernodes); the others run task lists: forward (supernode, 64 front rows),
        // backward (supernode, as many columns as groups of G = chol_bwd_group( r, wide ) lanes fit in
        // 256 threads).
        // A third list, (supernode, 64 update rows) over all supernodes, forms M = L_off inv(L_ss).
        std::vector<char> small( (size_t)nl, 0 ), small_bwd( (size_t)nl, 0 ), split( (size_t)nl, 0 );
        std::vector<CholTask> tasks;
        std::vector<int64_t> tptr[4];
        for( int kind = 0; kind < 4; ++kind ) { tptr[kind].assign( (size_t)nl + 1, 0 ); }
        auto task = [&]( int32_t s, int32_t k )
        {
            CholTask t;
            t.f = sn_begin[s]; t.w = sn_begin[s + 1] - sn_begin[s]; t.r = (int32_t)( row_ptr[s + 1] - row_ptr[s] ); t.k = k;
            t.rb = row_ptr[s]; t.vp = val_ptr[s];
            return t;
        };
        std::vector<CholTask> sdesc( (size_t)ns );
        for( size_t k = 0; k < lnodes.size(); ++k ) { sdesc[k] = task( lnodes[k], 0 ); }
        {
            std::vector<CholTask> tk[4];
            for( int32_t l = 0; l < nl; ++l )
            {
                small[(size_t)l] = rmax[(size_t)l] <= CHOL_SMALL_ROWS;
                // the backward's lane groups also serve levels of many slightly longer supernodes (one
                // task per supernode would launch too many wavefronts there)
                small_bwd[(size_t)l] = small[(size_t)l] || ( rmax[(size_t)l] <= CHOL_SMALL_ROWS_MANY &&
                                                             lptr[(size_t)l + 1] - lptr[(size_t)l] >= CHOL_MANY_SUPERNODES );
                for( int64_t k = lptr[(size_t)l]; k < lptr[(size_t)l + 1]; ++k )
                {
                    const int32_t s = lnodes[(size_t)k];
                    const int32_t w = sn_begin[s + 1] - sn_begin[s];
                    const int32_t r = (int32_t)( row_ptr[s + 1] - row_ptr[s] );
                    for( int32_t q0 = 0; q0 < std::max( r - w, 1 ); q0 += 64 ) { tk[2].push_back( task( s, q0 ) ); }
                    if( !small[(size_t)l] ) { for( int32_t a0 = 0; a0 < r; a0 += CHOL_FWD_ROWS ) { tk[0].push_back( task( s, a0 ) ); } }
                    if( !small_bwd[(size_t)l] )
                    {
                        const int per = 256 / chol_bwd_group( r, wmax[(size_t)l] > CHOL_NARROW_WIDTH );
                        for( int32_t c0 = 0; c0 < w; c0 += per ) { tk[1].push_back( task( s, c0 ) ); }
                    }
                }
                // wide levels of few forward tasks run the split forward: 16-row tasks after the front launch
                if( !small[(size_t)l] && wmax[(size_t)l] > CHOL_NARROW_WIDTH &&
                    (int64_t)tk[0].size() - tptr[0][(size_t)l] < CHOL_SPLIT_TASKS )
                {
                    split[(size_t)l] = 1;
                    for( int64_t k = lptr[(size_t)l]; k < lptr[(size_t)l + 1]; ++k )
                    {
                        const int32_t s = lnodes[(size_t)k];
                        const int32_t r = (int32_t)( row_ptr[s + 1] - row_ptr[s] );
                        for( int32_t a0 = 0; a0 < r; a0 += CHOL_SPLIT_ROWS ) { tk[3].push_back( task( s, a0 ) ); }
                    }
                }
                for( int kind = 0; kind < 4; ++kind ) { tptr[kind][(size_t)l + 1] = (int64_t)tk[kind].size(); }
            }
            int64_t off = 0;
            for( int kind = 0; kind < 4; ++kind )
            {
                for( auto & x : tptr[kind] ) { x += off; }
                tasks.insert( tasks.end(), tk[kind].begin(), tk[kind].end() );
                off = (int64_t)tasks.size();
            }
        }
        if( !tasks.empty() ) { ctx->h1_tasks.upload( tasks.data(), tasks.size(), ctx->stream ); }
        ctx->h1_sdesc.upload( sdesc.data(), sdesc.size(), ctx->stream );
        ctx->h1_sn_begin.upload( sn_begin, (size_t)ns + 1, ctx->stream );
        ctx->h1_row_ptr.upload( row_ptr, (size_t)ns + 1, ctx->stream );
        ctx->h1_rows.upload( rows, (size_t)nrows, ctx->stream );
        ctx->h1_val_ptr.upload( val_ptr, (size_t)ns + 1, ctx->stream );
        {
            // row-major copies of the blocks of every level whose backward runs lane groups (same offsets
            // as in h1_vals)
            int64_t extent = 1;
            for( int32_t l = 0; l < nl; ++l )
            {
                if( !small_bwd[(size_t)l] ) { continue; }
                for( int64_t k = lptr[(size_t)l]; k < lptr[(size_t)l + 1]; ++k ) { extent = std::max<int64_t>( extent, sdesc[(size_t)k].vp + (int64_t)sdesc[(size_t)k].w * sdesc[(size_t)k].r ); }
            }
            h1_load_values( ctx, vals, val_ptr[ns], tptr[2][0], tptr[2][(size_t)nl], lptr, small_bwd, extent );
            ctx->h1_nvals = val_ptr[ns];
            ctx->h1_vt_extent = extent;
        }
        ctx->h1_kid_ptr.upload( kid_ptr.data(), kid_ptr.size(), ctx->stream );
        ctx->h1_kids.upload( kids.data(), kids.size(), ctx->stream );
        ctx->h1_cmap.upload( cmap.data(), cmap.size(), ctx->stream );
        ctx->h1_pull2.upload( pull2.data(), pull2.size(), ctx->stream );
        ctx->h1_pull_src.upload( pull_src.data(), pull_src.size(), ctx->stream );
        ctx->h1_bp.require( (size_t)( 3 * (int64_t)n ) );
        ctx->h1_perm.upload( perm, (size_t)n, ctx->stream );
        {
            std::vector<int32_t> iperm( (size_t)n );
            for( int64_t j = 0; j < n; ++j ) { iperm[(size_t)perm[j]] = (int32_t)j; }
            ctx->h1_iperm.upload( iperm.data(), iperm.size(), ctx->stream );
            RP_HIP( hipStreamSynchronize( ctx->stream ) );
        }
        ctx->h1_level_nodes.upload( lnodes.data(), lnodes.size(), ctx->stream );
        ctx->h1_z.require( (size_t)(3 * nrows) );
        ctx->h1_fr.require( (size_t)(3 * nrows) );
        ctx->h1_xp.require( (size_t)(3 * (int64_t)n) );
        RP_HIP( hipStreamSynchronize( ctx->stream ) );
        ctx->h1_level_ptr = lptr;
        ctx->h1_level_wmax = wmax;
        ctx->h1_level_rmax = rmax;
        ctx->h1_level_g = gsz;
        ctx->h1_level_small = small;
        ctx->h1_level_small_bwd = small_bwd;
        for( int kind = 0; kind < 4; ++kind ) { ctx->h1_task_ptr[kind] = tptr[kind]; }
        ctx->h1_level_split = split;
        ctx->h1_n = n;
        ctx->h1_ns = ns;
        ctx->have_h1 = true;
        ctx->have_h1_structure = true;
        ++ctx->h1_version;
    });
}

int rp_h1_solve( rp_ctx * ctx, double alpha, const double * B, int64_t ldB, double beta, double * X, int64_t ldX,
                 int nrhs, int ptr_flags )
{
    return guarded( "rp_h1_solve", [&]()
    {
        require( ctx != nullptr, "ctx is NULL" );
        require_nrhs( nrhs );
        require( ctx->have_h1, "no H1 factor uploaded (call rp_upload_h1_factor)", RP_ENOTREADY );
        require( B != nullptr && X != nullptr && ldB >= 3 && ldX >= 3, "B and X must be V x 3 with ld >= 3" );
        RP_HIP( hipSetDevice( ctx->device ) );
        Staged st( ctx, ptr_flags & RP_DEVICE_PTR, B, ldB, beta, X, ldX );
        EventScope ev( ctx, 8 );
        h1_solve( ctx, alpha, st.X, st.ldx, beta, st.Y, st.ldy );
        ev.stop();
        RP_HIP( hipGetLastError() );
        st.finish();
        finish_timing( ctx, 8 );
    });
}

int rp_multiply_preconditioner( rp_ctx * ctx, double alpha, const double * X, int64_t ldX, double beta, double * Y, int64_t ldY,
                                int nrhs, int ptr_flags )
{
    return guarded( "rp_multiply_preconditioner", [&]()
    {
        require( ctx != nullptr, "ctx is NULL" );
        require_nrhs( nrhs );
        require( ctx->have_h1 && ctx->have_pl, "the preconditioner needs the H1 factor and the pseudo-Laplacian", RP_ENOTREADY );
        require( X != nullptr && Y != nullptr && ldX >= 3 && ldY >= 3, "X and Y must be V x 3 with ld >= 3" );
        RP_HIP( hipSetDevice( ctx->device ) );
        require( ctx->pc_weight_setting == 0.0 || ctx->have_metric, "the diagonal term needs the assembled metric", RP_ENOTREADY );
        Staged st( ctx, ptr_flags & RP_DEVICE_PTR, X, ldX, beta, Y, ldY );
        {
            EventScope ev( ctx, 10 );
            prepare_pc( ctx );
            ev.stop();
        }
        EventScope ev( ctx, 7 );
        pc_apply( ctx, alpha, st.X, st.ldx, beta, st.Y, st.ldy );
        ev.stop();
        RP_HIP( hipGetLastError() );
        st.finish();
        finish_timing( ctx, 7 );
    });
}

int rp_solve( rp_ctx * ctx, double alpha, const double * B, int64_t ldB, double beta, double * X, int64_t ldX,
              int nrhs, int max_iter, double tolerance, int solver_flags, int ptr_flags, int * iterations, double * rel_residuals )
{
    return guarded( "rp_solve", [&]()
    {
        require( ctx != nullptr, "ctx is NULL" );
        require_nrhs( nrhs );
        require( ctx->have_metric, "metric values not assembled (call rp_tp0_compute with RP_METRIC)", RP_ENOTREADY );
        const bool use_pc = !( solver_flags & RP_SOLVE_NO_PRECONDITIONER );
        require( !use_pc || (ctx->have_h1 && ctx->have_pl), "the preconditioner needs the H1 factor and the pseudo-Laplacian", RP_ENOTREADY );
        require( B != nullptr && X != nullptr && ldB >= 3 && ldX >= 3, "B and X must be V x 3 with ld >= 3" );
        require( max_iter >= 0 && tolerance >= 0, "max_iter and tolerance must be non-negative" );
        RP_HIP( hipSetDevice( ctx->device ) );
        const int64_t V = ctx->V;
        const size_t n3 = (size_t)(3 * V);
        Staged st( ctx, ptr_flags & RP_DEVICE_PTR, B, ldB, beta, X, ldX );
        ctx->cg_x.require( n3 ); ctx->cg_b.require( n3 );
        double * x = ctx->cg_x.get(); double * b = ctx->cg_b.get();
        if( use_pc )
        {
            EventScope ev( ctx, 10 );
            prepare_pc( ctx );
            ev.stop();
            RP_HIP( hipStreamSynchronize( ctx->stream ) );
            finish_timing( ctx, 10 );
        }
        const bool record = ( solver_flags & RP_SOLVE_RECORD ) && !( solver_flags & RP_SOLVE_GMRES );
        ctx->hist_n = 0;
        if( record )
        {
            ctx->hist_cap = std::min( max_iter, 1000 ) + 1;
            ctx->hist_r.require( (size_t)( 3 * V ) * ctx->hist_cap );
            ctx->hist_x.require( (size_t)( 3 * V ) * ctx->hist_cap );
        }
        EventScope ev( ctx, 3 );
        launch_copy3( V, st.X, st.ldx, b, 3, ctx->stream );
        std::vector<double> rel( 3, 0.0 );
        bool done = false;
        const int it = ( solver_flags & RP_SOLVE_GMRES )
                     ? gmres( ctx, b, x, max_iter, tolerance, use_pc, rel, done )
                     : pcg( ctx, b, x, max_iter, tolerance, use_pc, use_pc && ( solver_flags & RP_SOLVE_PC_NORM ), rel, done, record );
        launch_axpby3( V, alpha, x, 3, beta, st.Y, st.ldy, ctx->stream );
        ev.stop();
        RP_HIP( hipGetLastError() );
        st.finish();
        finish_timing( ctx, 3 );
        ctx->solve_iters = it;
        ctx->solve_relres = rel;
        if( iterations ) { *iterations = it; }
        if( rel_residuals ) { for( int k = 0; k < 3; ++k ) { rel_residuals[k] = rel[(size_t)k]; } }
        if( !done && ctx->guard_tripped )
        {
            throw InvalidArg{ "stopped by the residual guard at iteration " + std::to_string( it ) + " (relative residual above "
                              + std::to_string( ctx->guard_ratio ) + ")", RP_ESTAGNATION };
        }
        if( !done ) { throw InvalidArg{ "no convergence to the requested tolerance within max_iter iterations", RP_ESTAGNATION }; }
    });
}

int rp_set_solve_guard( rp_ctx * ctx, int check_iteration, double max_relative_residual )
{
    return guarded( "rp_set_solve_guard", [&]()
    {
        require( ctx != nullptr, "ctx is NULL" );
        require( check_iteration >= 0 && std::isfinite( max_relative_residual ) && max_relative_residual >= 0.0, "bad guard" );
        ctx->guard_iter = check_iteration;
        ctx->guard_ratio = max_relative_residual;
    });
}

int rp_set_preconditioner( rp_ctx * ctx, double diagonal_weight )
{
    return guarded( "rp_set_preconditioner", [&]()
    {
        require( ctx != nullptr, "ctx is NULL" );
        require( std::isfinite( diagonal_weight ), "the weight must be finite" );
        ctx->pc_weight_setting = diagonal_weight;
        ctx->pc_weight = diagonal_weight > 0 ? diagonal_weight : 0.0;
    });
}

int rp_preconditioner_weight( rp_ctx * ctx, double * weight )
{
    return guarded( "rp_preconditioner_weight", [&]()
    {
        require( ctx != nullptr && weight != nullptr, "bad arguments" );
        *weight = ctx->pc_weight;
    });
}

int rp_download_metric_diagonal( rp_ctx * ctx, double * diag )
{
    return guarded( "rp_download_metric_diagonal", [&]()
    {
        require( ctx != nullptr && diag != nullptr, "bad arguments" );
        require( ctx->have_metric, "metric values not assembled (call rp_tp0_compute with RP_METRIC)", RP_ENOTREADY );
        RP_HIP( hipSetDevice( ctx->device ) );
        ensure_metric_diag( ctx );
        ctx->dA.download( diag, (size_t)ctx->V, ctx->stream );
        RP_HIP( hipStreamSynchronize( ctx->stream ) );
    });
}

int rp_pc_set_model( rp_ctx * ctx, int nfeat, const float * features, const float * params, int bf16 )
{
    return guarded( "rp_pc_set_model", [&]()
    {
        require( ctx != nullptr && ctx->have_tree, "upload the tree first", RP_ENOTREADY );
        if( nfeat == 0 ) { ctx->lp_enabled = false; ctx->lp_scale_ready = false; return; }
        require( nfeat > 0 && nfeat <= MLP_FMAX && features && params, "bad model arguments" );
        RP_HIP( hipSetDevice( ctx->device ) );
        const int NP = mlp_param_count( nfeat );
        for( int q = 0; q <= NP; ++q ) { require( std::isfinite( params[q] ), "model parameters must be finite" ); }
        ctx->lp_F = nfeat;
        ctx->lp_feats.upload( features, (size_t)( ctx->C * nfeat ), ctx->stream );
        ctx->lp_params.upload( params, (size_t)NP, ctx->stream );
        ctx->lp_log_w = params[NP];
        ctx->lp_bf16 = bf16 != 0;
        RP_HIP( hipStreamSynchronize( ctx->stream ) );
        ctx->lp_enabled = true;
        ctx->lp_scale_ready = false;
    });
}

int rp_pc_download_scaling( rp_ctx * ctx, double * sqrt_d )
{
    return guarded( "rp_pc_download_scaling", [&]()
    {
        require( ctx != nullptr && sqrt_d != nullptr, "bad arguments" );
        require( ctx->lp_enabled, "no learned preconditioner set", RP_ENOTREADY );
        RP_HIP( hipSetDevice( ctx->device ) );
        if( !ctx->lp_scale_ready ) { learned_scales( ctx, ctx->lp_bf16 ); ctx->lp_scale_ready = true; }
        ctx->lp_sqrtd.download( sqrt_d, (size_t)ctx->V, ctx->stream );
        RP_HIP( hipStreamSynchronize( ctx->stream ) );
    });
}

int rp_pc_learn_step( rp_ctx * ctx, float lr, int step, float * params, float * m, float * v, double * loss, double * grad )
{
    return guarded( "rp_pc_learn_step", [&]()
    {
        require( ctx != nullptr && ctx->lp_F > 0, "set a learned preconditioner model first (rp_pc_set_model)", RP_ENOTREADY );
        require( ctx->have_metric && ctx->have_pl && ctx->have_h1, "metric, pseudo-Laplacian and H1 factor are needed", RP_ENOTREADY );
        require( ctx->hist_n >= 2, "record a PCG solve first (rp_solve with RP_SOLVE_RECORD)", RP_ENOTREADY );
        require( params && m && v && step >= 1, "bad arguments" );
        RP_HIP( hipSetDevice( ctx->device ) );
        const int F = ctx->lp_F;
        const int NP = mlp_param_count( F );
        const int64_t V = ctx->V, n3 = 3 * V;
        hipStream_t s = ctx->stream;
        // model -> FP32 vertex scalings (training uses the FP32 forward)
        ctx->lp_params.upload( params, (size_t)NP, s );
        ctx->lp_log_w = params[NP];
        learned_scales( ctx, false );
        ctx->lp_scale_ready = false;
        update_weight( ctx );
        const double * sd = ctx->lp_sqrtd.get();
        ctx->lp_acc.require( (size_t)( 7 * n3 + 3 * V ) );
        double * e = ctx->lp_acc.get(); double * y = e + n3; double * u = y + n3; double * Au = u + n3;
        double * sg = Au + n3; double * tv = sg + n3; double * Ja = tv + n3;
        double * gl = Ja + n3; double * gacc1 = gl + V; double * gacc2 = gacc1 + V;
        ctx->lp_t1.require( (size_t)n3 ); ctx->lp_t2.require( (size_t)std::max<int64_t>( n3, NP + 1 ) );
        double * a = ctx->lp_t1.get(); double * q = ctx->lp_t2.get();
        ctx->cg_scal.require( 32 ); ctx->cg_part.require( 3 * 1024 );
        double * sc = ctx->cg_scal.get(); double * part = ctx->cg_part.get();
        launch_fill( gacc1, 2 * V, 0.0, s );
        const int ns = ctx->hist_n - 1;
        const double * xfin = ctx->hist_x.get() + (size_t)n3 * ns;
        const double * rfin = ctx->hist_r.get() + (size_t)n3 * ns;
        auto dot = [&]( const double * X, const double * Y, int slot ) { launch_dot3( V, X, Y, part, sc + 3 * slot, s ); };
        double h_rr0[3];
        dot( ctx->hist_r.get(), ctx->hist_r.get(), 0 );
        RP_HIP( hipMemcpyAsync( h_rr0, sc, sizeof(h_rr0), hipMemcpyDeviceToHost, s ) );
        RP_HIP( hipStreamSynchronize( s ) );
        const double rr0 = h_rr0[0] + h_rr0[1] + h_rr0[2];
        // Per sample (r_k, e_k = x_last - x_k, y_k = r_k - r_last = A e_k, E_k = e_k . y_k) with u_k = P r_k:
        //   N = sum_k u_k.A u_k / E_k,  Cs = sum_k u_k.y_k / E_k,  loss = 1 - Cs^2 / (N K)
        // = min over alpha of mean_k |alpha u_k - e_k|_A^2 / |e_k|_A^2, invariant under scaling P. With the
        // optimal alpha = Cs/N held fixed, dloss/du_k = (alpha^2 g1_k - alpha g2_k) / K, g1_k = 2 A u_k / E_k,
        // g2_k = 2 y_k / E_k: the parameter gradients of g1 and g2 are accumulated apart and combined at the end.
        double Nsum = 0.0, Csum = 0.0, gw1 = 0.0, gw2 = 0.0;
        int used = 0;
        for( int k = 0; k < ns; ++k )
        {
            const double * r = ctx->hist_r.get() + (size_t)n3 * k;
            const double * xk = ctx->hist_x.get() + (size_t)n3 * k;
            launch_copy3( V, xfin, 3, e, 3, s );
            launch_axpby3( V, -1.0, xk, 3, 1.0, e, 3, s );              // e = x_last - x_k
            launch_copy3( V, r, 3, y, 3, s );
            launch_axpby3( V, -1.0, rfin, 3, 1.0, y, 3, s );             // y = r_k - r_last
            launch_row_scale( V, sd, r, 3, 1.0, 0.0, a, 3, s );          // a = D^{1/2} r
            pc2_apply( ctx, 1.0, a, 3, 0.0, q, 3 );                      // q = (P0 + w J) a
            launch_row_scale( V, sd, q, 3, 1.0, 0.0, u, 3, s );          // u = D^{1/2} q
            metric_apply( ctx, 1.0, u, 3, 0.0, Au, 3, true, true, false );
            double h[12];
            dot( e, y, 0 ); dot( u, Au, 1 ); dot( u, y, 2 ); dot( r, r, 3 );
            RP_HIP( hipMemcpyAsync( h, sc, sizeof(h), hipMemcpyDeviceToHost, s ) );
            RP_HIP( hipStreamSynchronize( s ) );
            const double Ek = h[0] + h[1] + h[2], Nk = h[3] + h[4] + h[5], Ck = h[6] + h[7] + h[8], rr = h[9] + h[10] + h[11];
            // e_k of the last iterates is dominated by the final error: use residuals down to 1e-6 |r_0|
            if( !( Ek > 0.0 ) || !( Nk > 0.0 ) || rr < 1e-12 * rr0 ) { continue; }
            ++used;
            Nsum += Nk / Ek; Csum += Ck / Ek;
            launch_fill( Ja, n3, 0.0, s );
            launch_jacobi_add( V, 1.0, a, 3, ctx->dA.get(), Ja, 3, s );  // J a
            for( int part_k = 0; part_k < 2; ++part_k )
            {
                double * g = part_k == 0 ? Au : y;                        // g1 = (2/E) A u, g2 = (2/E) y
                launch_scale1( n3, 2.0 / Ek, g, s );
                launch_row_scale( V, sd, g, 3, 1.0, 0.0, sg, 3, s );     // sg = D^{1/2} g
                pc2_apply( ctx, 1.0, sg, 3, 0.0, tv, 3 );                // t = (P0 + w J) D^{1/2} g
                launch_grad_logd( V, sd, g, q, tv, r, gl, s );
                launch_axpy1( V, 1.0, gl, part_k == 0 ? gacc1 : gacc2, s );
                double h_gw[3];
                dot( sg, Ja, 0 );
                RP_HIP( hipMemcpyAsync( h_gw, sc, sizeof(h_gw), hipMemcpyDeviceToHost, s ) );
                RP_HIP( hipStreamSynchronize( s ) );
                ( part_k == 0 ? gw1 : gw2 ) += ctx->pc_weight * ( h_gw[0] + h_gw[1] + h_gw[2] );
            }
        }
        require( used > 0 && Nsum > 0.0, "the recorded solve holds no usable residual/error pairs", RP_ENOTREADY );
        const double alpha = Csum / Nsum;
        const double loss_val = 1.0 - Csum * alpha / used;
        ctx->lp_alpha = alpha;
        launch_scale1( V, alpha * alpha, gacc1, s );
        launch_axpy1( V, -alpha, gacc2, gacc1, s );                      // gacc = alpha^2 gacc1 - alpha gacc2
        const double gw_sum = alpha * alpha * gw1 - alpha * gw2;
        const double inv = 1.0 / used;
        // dL/dlam per cluster: through the smoothing of log d (its adjoint), per primitive, then subtree
        // sums (leaf clusters, then levels)
        ctx->lp_gp.require( (size_t)ctx->T ); ctx->lp_gc.require( (size_t)ctx->C );
        launch_smooth_transpose( V, ctx->T, ctx->size, ctx->pre_cols.get(), ctx->vp_ptr.get(), ctx->vp_entry.get(), gacc1, ctx->lp_gp.get(), s );
        launch_grad_prim( ctx->T, ctx->size, ctx->pre_cols.get(), ctx->vp_ptr.get(), gacc1, ctx->lp_gp.get(), s );
        launch_up_leaves( ctx->L, ctx->leaf_cluster.get(), ctx->C_begin.get(), ctx->C_end.get(), ctx->lp_gp.get(), ctx->lp_gc.get(), s, 1 );
        launch_up_tree( ctx->h_level_ptr, ctx->level_nodes.get(), ctx->C_left.get(), ctx->C_right.get(), ctx->lp_gc.get(), s, 1 );
        launch_scale1( ctx->C, inv, ctx->lp_gc.get(), s );                                        // mean over samples
        ctx->lp_part.require( (size_t)( mlp_backward_blocks( ctx->C ) * NP ) );
        ctx->lp_grad.require( (size_t)( NP + 1 ) );
        launch_mlp_backward( ctx->C, F, ctx->lp_feats.get(), ctx->lp_params.get(), ctx->lp_gc.get(), ctx->lp_part.get(), ctx->lp_grad.get(), s );
        const double gw = gw_sum * inv;
        RP_HIP( hipMemcpyAsync( ctx->lp_grad.get() + NP, &gw, sizeof(double), hipMemcpyHostToDevice, s ) );
        if( lr > 0.0f )
        {
            ctx->lp_m.upload( m, (size_t)( NP + 1 ), s ); ctx->lp_v.upload( v, (size_t)( NP + 1 ), s );
            ctx->lp_t2.require( (size_t)( NP + 1 ) );
            float * p32 = reinterpret_cast<float *>( ctx->lp_t2.get() );     // all parameters incl. log w
            RP_HIP( hipMemcpyAsync( p32, params, sizeof(float) * ( NP + 1 ), hipMemcpyHostToDevice, s ) );
            launch_adam( NP + 1, p32, ctx->lp_m.get(), ctx->lp_v.get(), ctx->lp_grad.get(), lr, 0.9f, 0.999f, 1e-8f, step, s );
            RP_HIP( hipMemcpyAsync( params, p32, sizeof(float) * ( NP + 1 ), hipMemcpyDeviceToHost, s ) );
            ctx->lp_m.download( m, (size_t)( NP + 1 ), s ); ctx->lp_v.download( v, (size_t)( NP + 1 ), s );
        }
        if( grad ) { ctx->lp_grad.download( grad, (size_t)( NP + 1 ), s ); }
        RP_HIP( hipGetLastError() );
        RP_HIP( hipStreamSynchronize( s ) );
        if( lr > 0.0f )
        {
            ctx->lp_params.upload( params, (size_t)NP, s );
            ctx->lp_log_w = params[NP];
            RP_HIP( hipStreamSynchronize( s ) );
        }
        if( loss ) { *loss = loss_val; }
    });
}

} // extern "C"
