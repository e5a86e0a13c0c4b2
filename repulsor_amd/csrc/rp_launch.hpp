#pragma once
// Kernel argument blocks and host launchers shared by the translation units.

#include "rp_internal.hpp"

namespace rp
{
    struct NFArgs
    {
        int64_t L;
        int64_t T;
        const int32_t  * leaf_begin;
        const int32_t  * leaf_end;
        const int64_t  * col_ptr;
        const int32_t  * col_prim;
        const uint32_t * col_mask;
        const int64_t  * val_ptr;
        const double   * a;       // T
        const double   * x;       // 3 x T
        const double   * P;       // 6 x T
        const double   * rec;     // T x 10 records (a, x, P): the column gathers of the tangent-point kernels
        const double   * origin;  // 3 x L (mixed precision local origins)
        double           scale;   // 1/(n+1)
        int              size;    // n+1
        Expo             ex;
        double         * dx;      // 10 x T
        double         * epart;   // L
        void           * vals;    // 2 x slots (double or float)
        double         * diag;    // 2 x T
        float2         * vals32;  // FP64 metric only, may be NULL: FP32 copies of vals
        double         * diag32;  // with vals32: 2 x T, the diagonal summed from the copies
    };

    struct FFArgs
    {
        int64_t C;
        const int64_t * rp;
        const int32_t * ci;
        const double  * a;        // C
        const double  * x;        // 3 x C
        const double  * P;        // 6 x C
        Expo            ex;
        double        * dx;       // 10 x C
        double        * epart;    // C
        void          * vals;     // 2 x nnz
        double        * diag;     // 2 x C
    };

    void launch_nf( unsigned flags, int kind, bool mixed, const NFArgs & A, hipStream_t s );
    void launch_ff( unsigned flags, int kind, bool mixed, const FFArgs & A, hipStream_t s );
    constexpr int64_t REDUCE_BLOCKS = 1024;
    void launch_reduce_sum( const double * part, int64_t n, double * out, hipStream_t s );   // out: 1 + REDUCE_BLOCKS

    void launch_prep_prims( int64_t T, int nd, int size, const double * near, double * a, double * x, double * P, double * rec, hipStream_t s );
    void launch_prep_clusters( int64_t C, const double * far, double * a, double * x, double * P, hipStream_t s );
    void launch_leaf_origin( int64_t L, int64_t C, const int32_t * leaf_cluster, const double * cx, double * origin, hipStream_t s );
    void launch_leaf_far( int64_t L, int64_t C, const int32_t * leaf_anc, const int32_t * leaf_anc_len, const double * ff_dx, double * leaf_far, hipStream_t s );
    void launch_leaf_out( int64_t L, int64_t C, const int32_t * leaf_anc, const int32_t * leaf_anc_len, const double * C_out, double * leaf_out, hipStream_t s );
    void launch_hulls( int n, int64_t T, int64_t L, const int32_t * P_ordering, const int32_t * leaf_of_prim, const int32_t * S,
                       const double * V, const double * charges, const double * nf_dx, const double * vf_dh, const double * leaf_far, double * B, hipStream_t s );
    void launch_assemble( int64_t V, const int64_t * vc_ptr, const int32_t * vc_corner, const double * B, double alpha, double beta, double * Y, int64_t ldY, hipStream_t s );
    void launch_pre( int size, int64_t T, const int32_t * cols, const double * vals, const double * X, int64_t ldX, double * P_in, hipStream_t s );
    void launch_up_leaves( int64_t L, const int32_t * leaf_cluster, const int32_t * C_begin, const int32_t * C_end, const double * P_in, double * C_in, hipStream_t s, int W = 12 );
    void launch_up_level( int64_t n, const int32_t * nodes, const int32_t * C_left, const int32_t * C_right, double * C_in, hipStream_t s, int W = 12 );
    // the whole up-sweep (all depth levels, deepest first) in launches of three levels each
    void launch_up_tree( const std::vector<int32_t> & level_ptr, const int32_t * level_nodes, const int32_t * C_left, const int32_t * C_right,
                         double * C_in, hipStream_t s, int W );
    void launch_leaf_out3( int64_t L, int64_t C, const int32_t * leaf_anc, const int32_t * leaf_anc_len, const double * C_out, double * leaf_out, hipStream_t s );
    void launch_nf_spmm_post( bool mixed, int size, int64_t T, int64_t L, const int32_t * lb, const int32_t * le,
                              const int64_t * cp, const int32_t * cpr, const uint32_t * cm, const int64_t * vp,
                              const void * vals, const double * diag, const double * xin,
                              const double * leaf_out, int use_far, const double * pre_vals, double * S_out, hipStream_t s );
    void launch_ff_spmm( bool mixed, int64_t C, const int64_t * rp, const int32_t * ci, const void * vals, const double * diag, const double * xin, double * yout, hipStream_t s );
    void launch_post_prim( int size, int64_t T, const int32_t * leaf_idx_of_prim, const double * pre_vals, const double * leaf_out, int use_far, double * S_out, hipStream_t s );
    // perm non-null: row j of Y takes vertex perm[j] (Y in that numbering)
    void launch_post_vert( int64_t V, const int64_t * vp_ptr, const int32_t * vp_entry, const double * S_out, double alpha, double beta, double * Y, int64_t ldY, hipStream_t s,
                           const int32_t * perm = nullptr );
    void launch_map_index( int64_t n, const int32_t * map, const int32_t * in, int32_t * out, hipStream_t s );   // out[i] = map[in[i]]
    void launch_fill( double * p, int64_t n, double v, hipStream_t s );

    // pseudo-Laplacian (k_pl.hip)
    void launch_pl_assemble( int precision, const NFArgs & A, const FFArgs & B, bool do_nf, bool do_ff, hipStream_t s );
    void launch_pl_pre( int size, int64_t T, const int32_t * cols, const double * pre_vals, const double * X, int64_t ldX, double * P_in, hipStream_t s );
    void launch_pl_nf_spmm( int precision, int size, int64_t L, const int32_t * lb, const int32_t * le, const int64_t * cp,
                            const int32_t * cpr, const uint32_t * cm, const int64_t * vp, const void * vals, const double * diag,
                            const double * xin, const double * leaf_out, int use_far, const double * pre_vals, double * S_out, hipStream_t s );
    void launch_pl_post_prim( int size, int64_t T, const int32_t * leaf_idx_of_prim, const double * pre_vals, const double * leaf_out, int use_far, double * S_out, hipStream_t s );
    void launch_pl_ff_spmm( bool mixed, int64_t C, const int64_t * rp, const int32_t * ci, const void * vals, const double * diag, const double * xin, double * yout, hipStream_t s );

    // supernodal triangular solves (k_chol.hip)
    struct CholDev
    {
        const int32_t * sn_begin;   // ns + 1
        const int64_t * row_ptr;    // ns + 1
        const int32_t * rows;       // row sets, sorted, own columns first
        const int64_t * val_ptr;    // ns + 1
        const double  * vals;       // column-major r x w blocks: inv(L_ss) on top, M = L_off inv(L_ss) below
        const double  * vals_t;     // row-major copies of the blocks of the small levels (same offsets)
        const int32_t * kid_ptr;    // ns + 1
        const int32_t * kids;       // children in increasing order
        const int32_t * cmap;       // per row entry: position in the parent's row list
        const int32_t * perm;       // new -> old
        const int2    * pull2;      // per front row: its sources among the children's update rows (global
                                    // row indices, child order; -1 for none) when it has at most two, else
                                    // x = -2 - (first position in pull_src) and y = the end of its run there
        const int32_t * pull_src;   // source rows of the front rows with more than two, in child order
        const double  * bp;         // 3 x n: alpha B(perm)
        double        * z;          // 3 x rows: front vectors
        double        * xp;         // 3 x n: solution in the new numbering
    };
    constexpr int CHOL_MAX_WIDTH = 2048;  // columns per supernode
    constexpr int CHOL_NARROW_WIDTH = 128; // levels up to this width run the 4-wave forward (k_chol_fwd_big<4>)
    constexpr int CHOL_SMALL_ROWS = 64;   // levels whose supernodes have at most this many rows run a lane group per supernode
    constexpr int CHOL_SMALL_ROWS_MANY = 96;  // the backward also on levels of at most this many rows ...
    constexpr int CHOL_MANY_SUPERNODES = 16384;  // ... that hold at least this many supernodes
    constexpr int CHOL_SPLIT_TASKS = 64;  // wide levels with fewer 64-row forward tasks run the split forward
    constexpr int CHOL_SPLIT_ROWS = 16;   // rows per task of the split forward
    constexpr int CHOL_FWD_ROWS = 64;     // front rows per forward task (one per lane)
    constexpr int CHOL_BWD_RPL = 16;      // rows per lane the backward tasks of the narrow larger levels aim at
                                          // (16 against 8: 20 us less per preconditioner application, 32 and 4 slower)

    // Lanes per column in the backward tasks of the larger levels (a 256-thread task takes 256 / G
    // columns): on the wide levels min(64, r rounded up to a power of two); on the others the power of
    // two covering r / CHOL_BWD_RPL rows per lane (two batches of block-value loads), from 4 to 64.
    __host__ __device__ inline int chol_bwd_group( int r, bool wide )
    {
        const int need = wide ? r : ( r + CHOL_BWD_RPL - 1 ) / CHOL_BWD_RPL;
        int g = wide ? 1 : 4;
        while( g < need && g < 64 ) { g <<= 1; }
        return g;
    }
    void launch_chol_form_m( const CholTask * tasks, int64_t count, const double * Lin, double * Lout, hipStream_t s );
    void launch_chol_gather( int64_t n, const int32_t * perm, double alpha, const double * B, int64_t ldB, double * bp, hipStream_t s );
    void launch_chol_fwd_small( const CholDev & F, const CholTask * nodes, int64_t count, int G, int rmax, hipStream_t s );   // G lanes per supernode, G >= its width
    void launch_chol_fwd_big( const CholDev & F, const CholTask * tasks, int64_t count, int wmax, hipStream_t s );
    void launch_chol_fwd_split( const CholDev & F, const CholTask * front_tasks, int64_t front_count,
                                const CholTask * tasks, int64_t count, double * fr, hipStream_t s );
    void launch_chol_bwd_small( const CholDev & F, const CholTask * nodes, int64_t count, int G, int rmax, hipStream_t s );   // G lanes per supernode, G >= its width
    void launch_chol_bwd_big( const CholDev & F, const CholTask * tasks, int64_t count, int rmax, bool wide, hipStream_t s );
    void launch_chol_transpose( const CholTask * desc, int64_t count, const double * vals, double * vals_t, hipStream_t s );
    void launch_chol_scatter( int64_t n, const int32_t * perm, const double * xp, double beta, double * Y, int64_t ldY, hipStream_t s );
    void launch_chol_gather_scaled( int64_t n, const int32_t * perm, const double * sd, const double * B, int64_t ldB, double * bp, hipStream_t s );
    void launch_chol_scatter_scaled( int64_t n, const int32_t * perm, const double * xp, const double * sd, const double * X, int64_t ldX,
                                     double cw, const double * dA, double alpha, double beta, double * Y, int64_t ldY, hipStream_t s );

    // Krylov vector kernels (k_solve.hip)
    void launch_dot3( int64_t n, const double * u, const double * v, double * part, double * out, hipStream_t s );   // part: 3 x 1024
    void launch_cg_update( int64_t n, double * x, double * r, const double * p, const double * Ap, const double * rho,
                           const double * pAp, const double * active, double * part, double * rr, hipStream_t s );
    void launch_cg_dir( int64_t n, double * p, const double * z, const double * rho_new, const double * rho_old, const double * active, hipStream_t s );
    void launch_cg_check( const double * rr, const double * bb, double tol2, double * active, hipStream_t s );
    void launch_copy3( int64_t n, const double * src, int64_t lds, double * dst, int64_t ldd, hipStream_t s );
    void launch_axpby3( int64_t n, double alpha, const double * X, int64_t ldX, double beta, double * Y, int64_t ldY, hipStream_t s );
    // exact diagonal of the metric (k_diag.hip)
    struct DiagArgs
    {
        int64_t V, C;
        int size;
        int have_near, use_far;
        const int64_t  * vp_ptr;      // vertex -> (primitive, slot) entries
        const int32_t  * vp_entry;
        const double   * pre_vals;    // T x 4 x size
        const int32_t  * leaf_of_prim;
        const int32_t  * leaf_begin;
        const int32_t  * leaf_end;
        const int64_t  * col_ptr;
        const int32_t  * col_prim;
        const uint32_t * col_mask;
        const int64_t  * val_ptr;
        const double   * nf_diag;     // 2T
        const int32_t  * leaf_cluster;
        const int32_t  * C_parent;
        const int32_t  * C_begin;
        const int32_t  * C_end;
        const double   * ff_diag;     // 2C
        const double   * S;           // 2C root-path sums of ff_diag (set by launch_metric_diag)
        const double   * vf_diag;     // 2T very-near-field diagonal terms (NULL if none)
    };
    void launch_metric_diag( bool mixed, const DiagArgs & A, const void * nf_vals, double * S, double * dA, hipStream_t s );
    void launch_jacobi_add( int64_t V, double c, const double * X, int64_t ldX, const double * dA, double * Y, int64_t ldY, hipStream_t s );

    // per-state data from vertex coordinates (k_update.hip)
    void launch_near_far_rows( int n, int64_t T, const int32_t * P_ordering, const int32_t * S, const double * V, int64_t ldV,
                               const double * charges, double * near, double * far, hipStream_t s );
    void launch_cluster_leaves( int64_t L, const int32_t * leaf_cluster, const int32_t * C_begin, const int32_t * C_end,
                                const double * pfar, int singletons, double * C_far, hipStream_t s );
    void launch_cluster_level( int64_t n, const int32_t * nodes, const int32_t * C_left, const int32_t * C_right, double * C_far, hipStream_t s );

    // maximal safe step size (k_collide.hip)
    void launch_prim_stage( int n, int64_t T, const int32_t * P_ordering, const int32_t * S, const double * X, const double * U,
                            double t_max, double * box, int32_t * SV, double * PU, hipStream_t s );
    void launch_leaf_boxes( int64_t L, const int32_t * leaf_cluster, const int32_t * C_begin, const int32_t * C_end, const double * pbox,
                            double * cbox, hipStream_t s );
    void launch_level_boxes( int64_t n, const int32_t * nodes, const int32_t * C_left, const int32_t * C_right, double * cbox, hipStream_t s );
    void launch_traverse_step( int64_t n_in, const int2 * in, int2 * out, int64_t out_cap, int2 * leaves, int64_t leaf_cap,
                               unsigned long long * counters, const int32_t * C_left, const int32_t * C_right, const double * cbox, hipStream_t s );
    void launch_leaf_pair_candidates( int n, int64_t n_pairs, const int2 * pairs, const int32_t * C_begin, const int32_t * C_end,
                                      const int32_t * SV, const double * pbox, const double * cbox, int2 * cand, int64_t cap,
                                      unsigned long long * counter, hipStream_t s );
    void launch_pair_steps( int n, int64_t n_cand, const int2 * cand, const double * PU, double t_max, double eps, unsigned long long * result,
                            hipStream_t s );

    // very-near field (k_vf.hip)
    constexpr int VF_REC = 24;                          // doubles per pattern-entry record
    constexpr int64_t VF_LANE_EVALUATION_LIMIT = (int64_t)1 << 24;
    struct VFArgs
    {
        int64_t nnz, T;
        const int32_t * row;          // nnz: row primitive of each full-pattern entry (tree order)
        const int32_t * col;          // nnz: column primitive
        const int32_t * P_ordering;   // tree position -> simplex
        const int32_t * simplices;    // simplex vertices (caller's order)
        const double  * coords;       // V x 3
        const double  * a;            // T areas (tree order)
        const double  * P;            // 6 x T projectors
        double theta2;                // near separation parameter squared
        int max_ref;                  // deepest subdivision level
        Expo ex;
        int energy, diff, metric;     // tangent-point outputs requested
        double pl_sexp;               // pseudo-Laplacian exponent -n/2 - s
        int64_t cap;                  // evaluations per lane before the traversal stops with an error
        double * rec;                 // nnz x VF_REC
    };
    struct VFRowArgs
    {
        int64_t nrows, T;
        const int32_t * rows;         // rows with VF entries
        const int64_t * rp;           // T+1
        const int64_t * dg;           // T: first strictly upper entry
        const int64_t * tr;           // nnz: entry holding the metric values (itself or the transposed upper entry)
        const double  * rec;
        double * dx;                  // 10 x T (added), or NULL
        double * dh;                  // T x 3(n+1) (written on VF rows), with dx
        double * vals;                // tangent-point: 2 x nnz; pseudo-Laplacian: nnz
        double * diag;                // tangent-point: 2 x T; pseudo-Laplacian: T (written on VF rows)
        double * epart;               // nrows energy partials, or NULL
    };
    void launch_vf_pairs( int n, int mode, const VFArgs & A, hipStream_t s );
    void launch_vf_rows( int n, int mode, const VFRowArgs & A, hipStream_t s );
    void launch_vf_stats( int64_t nnz, const double * rec, double * out, hipStream_t s );
    void launch_vf_spmm_post( int size, int64_t nrows, const int32_t * rows, const int64_t * rp, const int32_t * ci, const double * vals,
                              const double * diag, const double * P_in, const double * pre_vals, double * S_out, hipStream_t s );
    void launch_vf_pl_spmm_post( int size, int64_t nrows, const int32_t * rows, const int64_t * rp, const int32_t * ci, const double * vals,
                                 const double * diag, const double * P_in, const double * pre_vals, double * S_out, hipStream_t s );

    // learned preconditioner (k_learn.hip)
    constexpr int MLP_H = 64;      // hidden width
    constexpr int MLP_FMAX = 16;   // maximum feature count
    inline int mlp_param_count( int F ) { return MLP_H * F + 2 * MLP_H + 1; }
    void launch_mlp_infer( bool bf16, int64_t C, int F, const float * feats, const float * p32, void * p16, double * lam, hipStream_t s );
    // Smoothing steps of log d (vertex mean of primitive means) after the leaf-to-root sums.
    constexpr int LP_SMOOTH_STEPS = 16;
    // sqrt_d = exp(log d / 2) with log d the smoothed vertex means of the leaf sums (logd: V, pm: T work).
    void launch_vertex_scale( int64_t V, int64_t T, int size, const int32_t * pre_cols, const int64_t * vp_ptr, const int32_t * vp_entry,
                              const int32_t * leaf_of_prim, const double * leaf_lam, double * logd, double * pm, double * sqrt_d, hipStream_t s );
    // g <- (S^T)^LP_SMOOTH_STEPS g (h: T work), the adjoint of the smoothing.
    void launch_smooth_transpose( int64_t V, int64_t T, int size, const int32_t * pre_cols, const int64_t * vp_ptr, const int32_t * vp_entry,
                                  double * g, double * h, hipStream_t s );
    void launch_row_scale( int64_t V, const double * sd, const double * X, int64_t ldX, double alpha, double beta, double * Y, int64_t ldY, hipStream_t s );
    void launch_grad_logd( int64_t V, const double * sd, const double * g, const double * q, const double * t, const double * r, double * out, hipStream_t s );
    void launch_grad_prim( int64_t T, int size, const int32_t * pre_cols, const int64_t * vp_ptr, const double * g_logd, double * g_prim, hipStream_t s );
    int64_t mlp_backward_blocks( int64_t C );
    void launch_mlp_backward( int64_t C, int F, const float * feats, const float * p, const double * g_lam, double * part, double * grad, hipStream_t s );
    void launch_adam( int NP, float * p, float * m, float * v, const double * grad, float lr, float b1, float b2, float eps, int t, hipStream_t s );
    // root-path sums of one value per cluster, per leaf (k_tree.hip)
    void launch_leaf_path_sum1( int64_t L, int64_t C, const int32_t * leaf_anc, const int32_t * leaf_anc_len, const double * src, double * leaf_out, hipStream_t s );

    // cluster tree and block cluster tree from coordinates (k_build.hip); the squared separation parameters
    void build_topology( BuiltTopology & B, int dom_dim, int64_t V, int64_t T, const double * V_coords, const int32_t * simplices,
                         int split_threshold, double far_theta2, double near_theta2, double inter_theta2, hipStream_t s );

    void launch_axpy1( int64_t n, double alpha, const double * x, double * y, hipStream_t s );
    void launch_scale1( int64_t n, double alpha, double * x, hipStream_t s );

    constexpr int MDOT_PART = 256 * 3;   // partial-sum doubles per basis vector of launch_mdot
    void launch_mdot( int64_t n, const double * Vb, int64_t stride, int nv, const double * w, double * part, double * out, int accumulate, hipStream_t s );
    void launch_maxpy( int64_t n, const double * Vb, int64_t stride, int nv, const double * coef, double sign, double * w, hipStream_t s );
    void launch_normalize3( int64_t n, const double * w, const double * nrm2, double * out, hipStream_t s );

} // namespace rp
