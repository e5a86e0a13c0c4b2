#pragma once
// Internal definitions shared by the HIP translation units of librepulsor_hip.so.

#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>
#include <memory>

#include "../../include/repulsor_hip.h"

namespace rp
{
    void set_error( const std::string & msg );

    struct HipError
    {
        int status;
        std::string msg;
    };

    inline void check_hip( hipError_t e, const char * what )
    {
        if( e != hipSuccess )
        {
            throw HipError{ (e == hipErrorOutOfMemory) ? RP_ENOMEM : RP_EHIP,
                            std::string(what) + ": " + hipGetErrorString(e) };
        }
    }

#define RP_HIP(x) ::rp::check_hip( (x), #x )

    // Device buffer owned by a context.
    template<typename T>
    struct DBuf
    {
        T * ptr = nullptr;
        size_t n = 0;

        DBuf() = default;
        DBuf( const DBuf & ) = delete;
        DBuf & operator=( const DBuf & ) = delete;
        ~DBuf() { release(); }

        void release()
        {
            if( ptr ) { (void)hipFree( ptr ); }
            ptr = nullptr; n = 0;
        }

        // Grows (never shrinks) to at least `count` elements; contents are not preserved on growth.
        void require( size_t count )
        {
            if( count <= n && ptr ) { return; }
            release();
            const size_t bytes = std::max<size_t>( 1, count ) * sizeof(T);
            RP_HIP( hipMalloc( (void**)&ptr, bytes ) );
            n = std::max<size_t>( 1, count );
        }

        void upload( const T * h, size_t count, hipStream_t s )
        {
            require( count );
            if( count ) { RP_HIP( hipMemcpyAsync( ptr, h, count * sizeof(T), hipMemcpyHostToDevice, s ) ); }
        }

        void download( T * h, size_t count, hipStream_t s ) const
        {
            if( count ) { RP_HIP( hipMemcpyAsync( h, ptr, count * sizeof(T), hipMemcpyDeviceToHost, s ) ); }
        }

        T * get() const { return ptr; }
    };

    // Exponent description shared by host and device.
    struct Expo
    {
        double q, p;
        double q_half, p_half;
        int    q_half_i, p_half_i;
        int    q_int, p_int;
        int    q_flag;            // 2: q/2 >= 1, 0: q == 0
        // metric kernel exponent s_exp = -((s-1) + n/2), s = (p-n)/q, with a fast path when 3*s_exp is an integer
        double s_exp;
        int    s_mode;            // 0: integer k, 1: k/3 via cbrt, 2: generic pow
        int    s_k;               // numerator
    };

    inline Expo make_expo( double q, double p, int n, double s_exp )
    {
        Expo e{};
        e.q = q; e.p = p;
        e.q_half = q / 2; e.p_half = p / 2;
        e.q_half_i = (int)e.q_half; e.p_half_i = (int)e.p_half;
        e.q_int = (e.q_half == (double)e.q_half_i);
        e.p_int = (e.p_half == (double)e.p_half_i);
        e.q_flag = (e.q_half >= 1) ? 2 : ((e.q_half == 0) ? 0 : -1);
        e.s_exp = s_exp;
        const double k3 = 3.0 * s_exp;
        const double kr = std::nearbyint( k3 );
        if( s_exp == std::nearbyint(s_exp) ) { e.s_mode = 0; e.s_k = (int)std::nearbyint(s_exp); }
        else if( std::abs( k3 - kr ) < 1e-12 ) { e.s_mode = 1; e.s_k = (int)kr; }
        else { e.s_mode = 2; e.s_k = 0; }
        (void)n;
        return e;
    }

    struct Timings
    {
        double ms[16] = {};
    };

    // One supernode's block of the H^1 factor, or one task on it: first column f, width w, rows r, first
    // front row (or column) k of the task, offsets of its row set and of its values.
    constexpr int NUM_XCD = 8;   // accelerator complex dies of an MI355X, each with its own L2

    struct CholTask
    {
        int32_t f, w, r, k;
        int64_t rb, vp;
    };

    // Cluster tree and block cluster tree built on the device (k_build.hip), kept until downloaded.
    struct BuiltTopology
    {
        int dom_dim = 0;
        int64_t T = 0, C = 0, L = 0;
        int depth = 0;
        // input and primitive data
        DBuf<double>  coords, pser, tser;          // 3V; (4 + 3(n+1)) T in the caller's order; same in tree order
        DBuf<int32_t> simp, ssimp, pre_cols;       // simplices, sorted simplices (caller's order), sorted simplices (tree order)
        // tree levels
        DBuf<int32_t> nb, ne, ndepth, nleft, nright, nnl, nj0, lvl_ind, lvl_incl;
        DBuf<int8_t>  nflag, ndir;
        DBuf<double>  nmid, nbox;
        DBuf<int32_t> ord, ord2, pnode, pnode2, gcnt, fb_list;
        DBuf<uint8_t> isl;
        DBuf<unsigned int> fb_ctr;
        std::vector<int64_t> level_ptr;
        const int32_t * final_ord = nullptr;       // points into ord or ord2
        DBuf<int32_t> desc, dleaf, g, lb;
        // tree (depth-first numbering)
        DBuf<int32_t> C_left, C_right, C_begin, C_end, C_depth, leaf_lookup, leaf_cluster, leaf_b, leaf_e;
        DBuf<double>  C_ser;
        // block cluster tree
        DBuf<unsigned long long> ctr;
        DBuf<int2>    front_a, front_b, far_pairs, blk_pairs;
        DBuf<uint64_t> keys_a, keys_b, far_up, blocks, vf_raw, inter_raw, vf, inter, nbr;
        int64_t n_far_up = 0, n_blocks = 0, n_vf = 0, n_inter = 0, n_nbr = 0, n_cols = 0, far_nnz = 0;
        DBuf<int64_t> nbr_rp, col_cnt, col_ptr, far_rp, nsel;
        DBuf<int32_t> col_prim, far_ci;
        DBuf<uint32_t> col_mask;
        DBuf<char>    tmp;                         // hipCUB temporary storage
    };

} // namespace rp

struct rp_ctx
{
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[16][2];
    hipEvent_t ev_check = nullptr;       // PCG: the residual read-back of an iteration has landed
    double * h_pinned = nullptr;         // 16 doubles of pinned host memory for the PCG read-backs

    // sizes
    int dom_dim = 0, size = 0, nd = 0;
    int64_t V = 0, T = 0, C = 0, L = 0;
    int max_depth = 0;
    int max_leaf = 0;

    // mesh (original order)
    rp::DBuf<int32_t> simplices;
    rp::DBuf<double>  charges;
    bool have_charges = false;
    rp::DBuf<double>  coords;
    // vertex -> corner CSR (corner = t*(n+1)+j, increasing) for the derivative assembler
    rp::DBuf<int64_t> vc_ptr;
    rp::DBuf<int32_t> vc_corner;

    // tree
    rp::DBuf<int32_t> C_left, C_right, C_begin, C_end, C_parent, P_ordering;
    // root-to-leaf paths: leaf_anc[d * L + l] = the cluster at depth d on leaf l's path (d = 0: the
    // root), leaf_anc_len[l] entries (the leaf's depth + 1)
    rp::DBuf<int32_t> leaf_anc, leaf_anc_len;
    rp::DBuf<int32_t> leaf_of_prim;      // tree position -> leaf cluster id
    rp::DBuf<int32_t> leaf_cluster;      // leaf index -> cluster id
    rp::DBuf<int32_t> pre_cols;          // T x (n+1)
    rp::DBuf<double>  pre_vals;          // T x 4 x (n+1)
    rp::DBuf<int64_t> vp_ptr;            // vertex -> (tree primitive, slot), sorted by primitive
    rp::DBuf<int32_t> vp_entry;          // primitive * (n+1) + slot
    std::vector<int32_t> h_level_ptr;    // internal clusters grouped by depth, deepest first
    rp::DBuf<int32_t> level_nodes;
    bool have_tree = false;

    // near tiles
    rp::DBuf<int32_t>  leaf_begin, leaf_end;
    rp::DBuf<int64_t>  col_ptr, val_ptr;
    rp::DBuf<int32_t>  col_prim;
    rp::DBuf<uint32_t> col_mask;
    int64_t nf_slots = 0;
    int64_t nf_cols = 0;
    bool have_near = false;

    // far pattern
    rp::DBuf<int64_t> far_rp;
    rp::DBuf<int32_t> far_ci;
    int64_t far_nnz = 0;
    bool have_far = false;

    // very-near field: full symmetric CSR over primitives (tree order)
    rp::DBuf<int64_t> vf_rp, vf_dg, vf_tr;   // T+1, T, nnz
    rp::DBuf<int32_t> vf_ci, vf_row, vf_rows; // nnz, nnz, rows with entries
    int64_t vf_nnz = 0, vf_nrows = 0;
    double vf_theta = 10.0;              // near separation parameter
    int vf_max_ref = 30;                 // deepest subdivision level of the tangent-point kernel
    rp::DBuf<double> vf_rec;             // nnz x VF_REC per-entry records
    rp::DBuf<double> vf_dh;              // T x 3(n+1) hull-vertex moment derivatives of the last differential
    rp::DBuf<double> vf_vals, vf_diag;   // 2 x nnz, 2T metric values
    rp::DBuf<double> pl_vf_vals, pl_vf_diag;   // nnz, T pseudo-Laplacian values
    bool have_vf_dh = false, have_vf_metric = false, have_pl_vf = false;
    double vf_stats[3] = { 0, 0, 0 };    // evaluations, deepest level, limit hit (last tangent-point compute)

    // maximal safe step size: swept boxes, traversal work lists, leaf pairs, counters
    rp::DBuf<double> col_pbox, col_cbox, col_U, col_PU;
    rp::DBuf<int32_t> col_SV;
    rp::DBuf<int2> col_list_a, col_list_b, col_leaves, col_cand;
    rp::DBuf<unsigned long long> col_counters;

    // data (tree order, SoA)
    rp::DBuf<double> p_a, p_x, p_P;      // T, 3T, 6T
    rp::DBuf<double> p_rec;              // T x 10 records (a, x, P)
    rp::DBuf<double> c_a, c_x, c_P;      // C, 3C, 6C
    rp::DBuf<double> leaf_origin;        // 3L
    rp::DBuf<double> stage;              // upload staging; after rp_update_coordinates: the near rows T x ND
    rp::DBuf<double> upd_far;            // after rp_update_coordinates: primitive far rows T x 10, cluster rows C x 10
    bool have_state_rows = false;
    bool have_data = false;

    // energy/derivative outputs
    rp::DBuf<double> nf_dx;              // 10 x T
    rp::DBuf<double> ff_dx;              // 10 x C
    rp::DBuf<double> leaf_far;           // 10 x L: percolated far derivatives per leaf
    rp::DBuf<double> partials;           // L + C energy partials
    rp::DBuf<double> scalar;             // device scalars
    rp::DBuf<double> hull_grad;          // T x (n+1) x 3
    bool have_diff = false;

    // metric
    rp::DBuf<double> nf_vals;            // 2 x nf_slots (fp64) or packed floats
    rp::DBuf<double> nf_diag;            // 2T
    rp::DBuf<double> nf_vals32;          // RP_MIXED_REFINED: nf_slots float2 copies of nf_vals
    rp::DBuf<double> nf_diag32;          // RP_MIXED_REFINED: 2T, the diagonal summed from the copies
    rp::DBuf<double> ff_vals;            // 2 x far_nnz
    rp::DBuf<double> ff_diag;            // 2C
    bool have_metric = false;
    int metric_precision = 0;

    // multiply buffers
    rp::DBuf<double> P_in, P_out, C_in, C_out, leaf_out;   // 12T, 12T, 12C, 12C, 12L
    rp::DBuf<double> Xd, Yd;                                // 3V staging

    // pseudo-Laplacian metric values (scalar blocks)
    rp::DBuf<double> pl_nf_vals, pl_nf_diag, pl_ff_vals, pl_ff_diag;
    bool have_pl = false;
    int pl_precision = 0;
    double pl_s = 0.0;

    // supernodal Cholesky factor of the H^1 metric
    rp::DBuf<int32_t> h1_sn_begin, h1_rows, h1_kid_ptr, h1_kids, h1_cmap, h1_perm, h1_level_nodes;
    rp::DBuf<int64_t> h1_row_ptr, h1_val_ptr;
    rp::DBuf<double>  h1_vals, h1_z, h1_xp, h1_bp, h1_fr;  // h1_vals: inv(L_ss) over M = L_off inv(L_ss) per supernode
    rp::DBuf<double>  h1_vals_t;                     // row-major copies of the small levels' blocks
    rp::DBuf<int32_t> h1_pull2, h1_pull_src;         // per front row: the child update rows that map to it
    std::vector<int64_t> h1_level_ptr;   // supernodes grouped by height (leaves first)
    std::vector<int>     h1_level_wmax;
    std::vector<int>     h1_level_rmax;
    std::vector<int>     h1_level_g;       // small levels: lanes per supernode
    std::vector<char>    h1_level_small;           // forward: the level runs a lane group per supernode
    std::vector<char>    h1_level_small_bwd;       // backward: the same
    rp::DBuf<rp::CholTask> h1_tasks;               // forward, backward and M-forming tasks
    rp::DBuf<rp::CholTask> h1_sdesc;               // per supernode, in level order
    std::vector<int64_t> h1_task_ptr[4];           // per kind (forward, backward, forming M, split forward): level offsets into h1_tasks
    std::vector<char>    h1_level_split;           // the forward of the level runs in two launches (front, then 16-row tasks)
    int64_t h1_n = 0, h1_ns = 0;
    int64_t h1_nvals = 0, h1_vt_extent = 1;   // stored factor values; slots of the row-major copies
    bool have_h1 = false;             // factor values of the current coordinates
    bool have_h1_structure = false;   // the structure of the last uploaded factor (kept while the topology is)

    // Krylov solver state
    rp::DBuf<double> cg_x, cg_r, cg_p, cg_Ap, cg_z, cg_b, pc_tmp, cg_scal, cg_part;
    rp::DBuf<double> gm_V, gm_w, gm_t, gm_small, gm_part;   // GMRES basis (m+1) x 3V and work

    // two-level preconditioner P0 + w diag(A)^-1: exact metric diagonal and the weight
    rp::DBuf<double> dA, dA_path;          // V, 2C
    int64_t metric_version = 0, pl_version = 0, h1_version = 0, tree_version = 0;
    // inside the preconditioner the H^1 solves hand the pseudo-Laplacian their vectors in the factor's
    // numbering: pre_cols mapped through the inverse permutation, rebuilt when the tree or factor changes
    rp::DBuf<int32_t> h1_iperm, pre_cols_h1;
    int64_t pre_cols_h1_key[2] = { -1, -1 };
    int64_t diag_version = -1;
    double pc_weight_setting = 0.0;        // 0: P0 only, > 0: fixed weight, < 0: estimated per state
    double pc_weight = 0.0;                // weight in use
    double pc_weight_est = 0.0;            // cached estimate lambda_max(P0 A) / 4
    int64_t pc_weight_key[3] = { -1, -1, -1 };

    // learned preconditioner D^{1/2} (P0 + w diag(A)^-1) D^{1/2}
    bool lp_enabled = false;
    bool lp_bf16 = true;
    int lp_F = 0;
    rp::DBuf<float>  lp_feats;             // C x F node features
    rp::DBuf<float>  lp_params;            // MLP parameters (FP32)
    rp::DBuf<float>  lp_params16;          // BF16 copy (as raw storage)
    rp::DBuf<float>  lp_m, lp_v;           // Adam moments
    double lp_log_w = 0.0;                 // learned log factor on the base diagonal weight
    double lp_alpha = 1.0;                 // optimal global scale of P from the last training step
    bool lp_scale_ready = false;
    rp::DBuf<double> lp_lam, lp_leaf, lp_sqrtd, lp_logd;     // C, L, V, V
    rp::DBuf<double> lp_t1, lp_t2;                   // 3V work
    rp::DBuf<double> lp_gl, lp_gp, lp_gc, lp_part, lp_grad, lp_acc;   // training
    // PCG guard (rp_set_solve_guard): at iteration guard_iter (> 0) a solve whose largest relative residual is
    // above guard_ratio stops with RP_ESTAGNATION; guard_tripped tells the caller why it stopped
    int guard_iter = 0;
    double guard_ratio = 0.0;
    bool guard_tripped = false;
    // PCG history for training pairs: r_k and x_k, k = 0..hist_n-1 (3V each)
    rp::DBuf<double> hist_r, hist_x;
    int hist_n = 0, hist_cap = 0;
    int solve_iters = 0;
    std::vector<double> solve_relres;

    rp::Timings timings;

    // device-built topology awaiting download (rp_build_topology)
    std::unique_ptr<rp::BuiltTopology> built;
};
