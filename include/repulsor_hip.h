/* C-ABI of the MI355X tangent-point library (librepulsor_hip.so).
 *
 * This is the drop-in boundary at the traversor seam of HenrikSchumacher/Repulsor: the header-only
 * C++ host code (include/repulsor/) keeps building the cluster tree, the block cluster tree and the
 * per-primitive data on the host, and hands them to these entry points, which run the
 * TangentPointEnergy0 near/far kernels, derivative collection, TangentPointMetric0's
 * MultiplyMetric and Solve on one AMD Instinct GPU (gfx950).
 *
 * Conventions: plain pointers and sizes only; all arrays row-major; int status return
 * (0 = RP_OK); rp_last_error() gives a thread-local message. Host pointers unless the
 * RP_DEVICE_PTR flag is passed. Calls are synchronous on return. A context is not reentrant.
 */
#ifndef REPULSOR_HIP_H
#define REPULSOR_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RP_OK            0
#define RP_EINVAL        1
#define RP_EHIP          2
#define RP_ENOMEM        3
#define RP_ESTAGNATION   4
#define RP_EUNSUPPORTED  5
#define RP_ENOTREADY     6

/* rp_tp0_compute flags */
#define RP_ENERGY        1u
#define RP_DIFF          2u
#define RP_METRIC        4u

/* pointer-location flag */
#define RP_DEVICE_PTR    1

/* precision */
#define RP_FP64          0
#define RP_MIXED         1
/* RP_MIXED_REFINED (metric only): FP64 pair arithmetic and FP64 metric values as RP_FP64, plus FP32
 * copies of the near-field metric values with a diagonal summed from those copies. rp_multiply_metric
 * uses the FP64 values; the PCG iterations of rp_solve use the FP32 copies until their recursive
 * residual meets the tolerance, then replace the residual by the FP64 one and, where that is still
 * above the tolerance, continue on the FP64 values. No reference counterpart. */
#define RP_MIXED_REFINED 2
/* RP_PL_BF16 (rp_pl_compute only; no reference counterpart): the pseudo-Laplacian inside the
 * preconditioner with its near-field values computed in FP32 and stored in BF16 (the diagonal summed
 * from the rounded values, so constants stay in its kernel), the far field in FP32. The preconditioner
 * may be inexact; the solve is judged by the metric's residual. */
#define RP_PL_BF16       3

typedef struct rp_ctx rp_ctx;

/* Thread-local message of the last failing call. */
const char * rp_last_error(void);

/* Library version string. */
const char * rp_version(void);

/* Context on HIP device `device` with its own stream (default priority).
 * Replaces: the mesh's persistent cache entries that own ClusterTree/BlockClusterTree
 * (reference src/SimplicialMesh.hpp:470-607). */
int rp_ctx_create( int device, rp_ctx ** out );
int rp_ctx_destroy( rp_ctx * ctx );

/* Mesh topology: simplices T x (dom_dim+1) in the caller's order, optional vertex charges (V).
 * Replaces: the SimplicialMesh pointer constructor (reference src/SimplicialMesh.hpp:143-185). */
int rp_upload_mesh( rp_ctx * ctx, int dom_dim, int64_t vertex_count, int64_t simplex_count,
                    const int32_t * simplices, const double * vertex_charges );

/* Cluster tree arrays (depth-first numbering, C clusters), P_ordering (tree position -> simplex),
 * and the mixed pre/post operator: pre_cols T x (n+1) vertex ids, pre_vals T x 4 x (n+1).
 * Replaces: ClusterTree construction outputs C_left/C_right/C_begin/C_end/P_ordering and
 * mi_pre/mi_post (reference src/ClusterTree.hpp:115-180, 410-639). */
int rp_upload_tree( rp_ctx * ctx, int32_t cluster_count,
                    const int32_t * C_left, const int32_t * C_right,
                    const int32_t * C_begin, const int32_t * C_end,
                    const int32_t * P_ordering,
                    const int32_t * pre_cols, const double * pre_vals );

/* Near field as leaf tiles: per leaf I the primitive range [leaf_begin, leaf_end), the
 * concatenated column primitives of all leaves sharing a non-admissible block with I
 * (col_ptr has leaf_count+1 entries), and a row-bit mask per column (bit ii: pair
 * (leaf_begin+ii, column) is a Near pair). Leaves hold at most 32 primitives.
 * Replaces: BlockClusterTree::Near() pattern (reference src/BlockClusterTree.hpp:234-330). */
int rp_upload_near_tiles( rp_ctx * ctx, int32_t leaf_count,
                          const int32_t * leaf_begin, const int32_t * leaf_end,
                          const int64_t * col_ptr, const int32_t * col_prim, const uint32_t * col_mask );

/* Far field: full symmetric CSR over clusters (rows = cluster_count).
 * Replaces: BlockClusterTree::Far() (reference src/BlockClusterTree.hpp:321-324). */
int rp_upload_far_pattern( rp_ctx * ctx, int32_t rows, const int64_t * rp, const int32_t * ci );

/* Very-near field: full symmetric CSR over primitives (rows = simplex count, tree order, no diagonal)
 * of the pairs that are neither adjacent nor separated by the near-field MAC, with the near separation
 * parameter theta and the deepest subdivision level of the tangent-point kernel (clamped as the
 * reference clamps it to the hierarchy's maximal level). rp[rows] = 0 clears it.
 * Replaces: BlockClusterTree::VeryNear() (reference src/BlockClusterTree.hpp:166, built at :313) and
 * AdaptivitySettings{max_refinement, theta} (src/Settings.hpp:49-54). */
int rp_upload_verynear_pattern( rp_ctx * ctx, int32_t rows, const int64_t * rp, const int32_t * ci,
                                double theta, int max_refinement );

/* Builds the cluster tree and the block cluster tree on the device from the clustering coordinates
 * V_coords (V x 3 row-major) and the simplices (T x (n+1)): longest-axis splits down to split_threshold
 * (1..32) primitives per leaf, depth-first numbering, the far field (AABB MAC with far_theta), the leaf
 * blocks and, inside them, the very-near and intersecting primitive pairs (primitive MAC with near_theta,
 * intersection test with intersection_theta). The results stay on the device until
 * rp_download_built_topology and equal the host construction of include/repulsor/ClusterTree.hpp and
 * BlockClusterTree.hpp element for element. The median-split fallback sorts the range stably by the
 * interior coordinate (one of the orders std::nth_element may leave); a fallback on a range of more than
 * 2048 primitives returns RP_EUNSUPPORTED (build on the host then). counts (8): clusters, leaves, depth,
 * far nnz (full symmetric), leaf blocks, very-near pairs, intersecting pairs, near-tile columns.
 * Replaces: ClusterTree construction (reference src/ClusterTree.hpp:115-180, ClusterTree/Split_Thread.hpp:3-166,
 * ClusterTree/Serialize_Thread.hpp:3-156, BoundingVolumes/AABB_LongestAxisSplit.hpp:134-253) and
 * BlockClusterTree::ComputeBlocks (src/BlockClusterTree.hpp:234-330, ClusterTreePairTraversor/Traverse_DepthFirst.hpp:3-267,
 * BlockClusterTree/BlockSplit_Kernel.hpp:164-251). */
int rp_build_topology( rp_ctx * ctx, int dom_dim, int64_t vertex_count, int64_t simplex_count, const double * V_coords,
                       const int32_t * simplices, int split_threshold, double far_theta, double near_theta, double intersection_theta,
                       int64_t * counts );

/* Copies the device-built topology to host arrays (any pointer may be NULL) and, with release != 0,
 * frees its device memory. Tree: P_ordering (T), C_left, C_right, C_begin, C_end, C_depth (C),
 * C_serialized (C x 7: r^2, box center, half extents). Block cluster tree: far_rp (C + 1), far_ci
 * (far nnz; full symmetric, sorted, no diagonal), leaf_blocks (2 x blocks: (I, J), I <= J, sorted),
 * verynear_pairs and inter_pairs (2 x count: tree-ordered primitive pairs (i, j), i < j, sorted), near
 * tiles col_ptr (leaves + 1), col_prim and col_mask (tile columns) in the rp_upload_near_tiles layout.
 * Replaces: the reference's accessors of the built trees (ClusterTree::PrimitiveOrdering, ClusterLeft, ...,
 * BlockClusterTree::Near/Far/VeryNear, src/ClusterTreeBase.hpp, src/BlockClusterTreeBase.hpp:25-28). */
int rp_download_built_topology( rp_ctx * ctx, int32_t * P_ordering, int32_t * C_left, int32_t * C_right, int32_t * C_begin,
                                int32_t * C_end, int32_t * C_depth, double * C_serialized, int64_t * far_rp, int32_t * far_ci,
                                int32_t * leaf_blocks, int32_t * verynear_pairs, int32_t * inter_pairs, int64_t * col_ptr,
                                int32_t * col_prim, uint32_t * col_mask, int release );

/* Per-state data, after construction and after every SemiStaticUpdate:
 * V_coords V x 3, P_near T x (1+3(n+1)+6) and C_far C x 10 in tree order.
 * Replaces: ClusterTree::SemiStaticUpdate data path (reference src/ClusterTree.hpp:983-993). */
int rp_upload_data( rp_ctx * ctx, const double * V_coords, const double * P_near, const double * C_far );

/* Per-state data computed on the device from new vertex coordinates (V x 3, leading dimension ldV,
 * host pointer or RP_DEVICE_PTR): simplex rows (area, hull, barycenter, projector) in tree order and
 * the cluster moments (area-weighted averages, leaves then levels up to the root). The results equal,
 * bit for bit, rp_upload_data of the host's rows for the same coordinates (the same operations in the
 * same order, no contraction). singletons != 0: leaves copy their primitive's row (split threshold
 * <= 1). Tree, patterns and pre/post operators stay as uploaded. Device-pointer coordinates are not
 * checked for NaN.
 * Replaces: SimplicialMesh::SemiStaticUpdate (reference src/SimplicialMesh.hpp:354-376) with
 * ComputeNearFarData (src/SimplicialMesh/ComputeNearFarData.hpp:3-27) and ClusterTree::ComputeClusterData
 * (src/ClusterTree/ComputeClusterData_Thread.hpp:50-129). */
int rp_update_coordinates( rp_ctx * ctx, const double * V_coords, int64_t ldV, int singletons, int ptr_flags );

/* Maximal step t in [0, max_time] such that moving the vertices by t U (U: V x 3 update vectors with
 * leading dimension ldU, host memory or RP_DEVICE_PTR) causes no collision between simplices that share
 * no vertex, as the reference computes it: per simplex pair, bisection of [0, max_time] until the
 * interval is shorter than TOL times its end, testing whether the space-time convex hulls of the two
 * moving simplices over the interval meet (GJK in R^4 with the reference's tolerance); pairs whose
 * boxes over [0, max_time] do not meet are skipped after a traversal of the cluster tree. Every pair
 * starts at max_time (the reference starts later pairs at its running minimum), so the result is
 * independent of the order of evaluation. stats (may be NULL): leaf pairs, simplex pairs tested, hull
 * tests. Uses the coordinates of the current state (also after rp_update_coordinates).
 * Replaces: SimplicialMesh::MaximumSafeStepSize (reference src/SimplicialMesh.hpp:441-468),
 * CollisionTree::MaximumSafeStepSize (src/CollisionTree.hpp), CollisionFinder::FindMaximumSafeStepSize
 * (src/CollisionFinder.hpp). */
int rp_max_safe_step( rp_ctx * ctx, const double * U, int64_t ldU, double max_time, double TOL, int ptr_flags,
                      double * t_out, int64_t * stats );

/* After rp_update_coordinates: the device's near rows (T x ND, tree order), cluster rows (C x 10) and
 * vertex coordinates (V x 3). Any pointer may be NULL. No reference counterpart (the host mesh reads
 * its state back through it when a host computation needs the coordinates). */
int rp_download_state_rows( rp_ctx * ctx, double * P_near, double * C_far, double * V_coords );

/* TangentPointEnergy0 kernels over the very-near, near and far fields (flags: RP_ENERGY | RP_DIFF |
 * RP_METRIC). energy_out may be NULL. RP_METRIC stores the metric blocks on the device for
 * rp_multiply_metric. precision: RP_FP64, RP_MIXED, or RP_MIXED_REFINED (computes like RP_FP64 and
 * with RP_METRIC also keeps the FP32 copies). The very-near field is computed in FP64 at every precision; its energy terms
 * are a E b per adaptive evaluation without the quadrature weight, as the reference sums them
 * (TP0_Kernel_VF.hpp:273), while its derivative and metric terms carry the weight. A very-near
 * traversal that exceeds 2^24 evaluations on one level-1 sub-pair stops with RP_EUNSUPPORTED.
 * Replaces: TP0_Traversor<...>::Compute (reference src/Energies/TangentPointEnergy0/TP0_Traversor.hpp:129-253). */
int rp_tp0_compute( rp_ctx * ctx, unsigned flags, double q, double p, int precision, double * energy_out );

/* Doubles of device scratch that rp_tp0_energy_noalloc can use (16 MiB per device). */
#define RP_NOALLOC_SCRATCH_DOUBLES (2 * 1024 * 1024)

/* No counterpart in the reference. For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): the energy of rp_tp0_compute(ctx,
 * RP_ENERGY, q, p, precision, energy_out), bit for bit, computed by the same kernels. It calls no hipMalloc or
 * hipFree: its energy partials (one per leaf, cluster and very-near row), reduction blocks and very-near
 * per-entry records (24 doubles per entry) live in a static device array of RP_NOALLOC_SCRATCH_DOUBLES
 * doubles that each device holds until the process ends. A mesh that needs more returns RP_EINVAL with the
 * size in the message (1.31M triangles with split_threshold 16 need about 0.34M doubles). All calls on
 * one device share that array: calls that overlap on one device race. energy_out is a host pointer, or a
 * device pointer with RP_DEVICE_PTR. The context must hold the uploaded mesh, tree, patterns and data. */
int rp_tp0_energy_noalloc( rp_ctx * ctx, double q, double p, int precision, double * energy_out, int ptr_flags );

/* Moment derivatives of the last RP_DIFF compute: P_D_near T x ND, P_D_far T x 10, original order.
 * Replaces: ClusterTree::CollectDerivatives (reference src/ClusterTree.hpp:1115-1208). */
int rp_download_derivatives( rp_ctx * ctx, double * P_D_near, double * P_D_far );

/* Y = alpha * dE/dV + beta * Y for the last RP_DIFF compute (V x 3, leading dimension ldY).
 * Replaces: SimplicialMesh::Assemble_ClusterTree_Derivatives (reference src/SimplicialMesh.hpp:651-675). */
int rp_assemble_vertex_gradient( rp_ctx * ctx, double alpha, double beta, double * Y, int64_t ldY, int ptr_flags );

/* Metric blocks of the last RP_METRIC compute in tile layout (nf: 2 x slots, ordered [leaf][column][row]),
 * NF diagonal T x 2, FF values nnz x 2 (CSR order of the far pattern), FF diagonal C x 2.
 * Any pointer may be NULL. Replaces: MetricValueContainer read-out (reference src/FMM/FMM_Configurator.hpp:5-20). */
int rp_download_metric( rp_ctx * ctx, double * nf_vals, double * nf_diag, double * ff_vals, double * ff_diag );

/* Very-near-field metric values of the last RP_METRIC compute: vf_vals nnz x 2 (a0, a1) in the CSR
 * order of the VF pattern, vf_diag T x 2 (tree order). The values are the reference kernel's per-pair
 * sums (TP0_Kernel_VF::compute, src/Energies/TangentPointEnergy0/TP0_Kernel_VF.hpp:300-331); the
 * reference stores them with a stride of 7 (off-diagonal) and 16 (diagonal) but reads them with a
 * stride of 2 (TP0_Traversor.hpp:453-478), which this library does not reproduce. Either pointer may
 * be NULL. */
int rp_download_vf_metric( rp_ctx * ctx, double * vf_vals, double * vf_diag );

/* Statistics of the very-near field of the last rp_tp0_compute: adaptive evaluations and deepest
 * subdivision level (FMM_Kernel_VF::PrintReport, src/FMM/FMM_Kernel_VF_Common.hpp:119-133). */
int rp_vf_stats( rp_ctx * ctx, int64_t * evaluations, int * max_level );

/* Y = alpha * A X + beta * Y, X and Y V x nrhs (nrhs must be 3), ld >= 3.
 * nf_flag = RP_NF_FP32_COPIES (metric assembled at RP_MIXED_REFINED): the near field multiplies with
 * the FP32 copies, the product the CG iterations of rp_solve use (no reference counterpart).
 * Replaces: MetricDimRestricted::MultiplyMetric (reference src/Metrics/MetricDimRestricted.hpp:109-139). */
#define RP_NF_FP32_COPIES 2
int rp_multiply_metric( rp_ctx * ctx, double alpha, const double * X, int64_t ldX,
                        double beta, double * Y, int64_t ldY, int nrhs,
                        int vf_flag, int nf_flag, int ff_flag, int ptr_flags );

/* Pseudo-Laplacian L_s (scalar blocks -(r^2)^(-n/2-s) on the same near/far patterns, diagonal
 * -sum_j (a_j/a_i) val_ij), assembled on the device from the uploaded data.
 * Replaces: PseudoLaplacian<Mesh,false>::compute_metric (reference src/Metrics/PseudoLaplacian.hpp:54-64,
 * kernels src/Metrics/PseudoLaplacian/PseudoLaplacian_Kernel_NF.hpp:124-146, _FF.hpp:101-146). */
int rp_pl_compute( rp_ctx * ctx, double s, int precision );

/* Read-out of the pseudo-Laplacian values (near values in tile order, far values in CSR order).
 * Any pointer may be NULL. Replaces: the PseudoLaplacian MetricValues container read-out. */
int rp_download_pl( rp_ctx * ctx, double * nf_vals, double * nf_diag, double * ff_vals, double * ff_diag );

/* Very-near-field pseudo-Laplacian values of the last rp_pl_compute: vf_vals nnz (CSR order of the VF
 * pattern), vf_diag T. Replaces: PseudoLaplacian_Traversor::VF_Compute
 * (reference src/Metrics/PseudoLaplacian/PseudoLaplacian_Traversor.hpp:176-202). */
int rp_download_pl_vf( rp_ctx * ctx, double * vf_vals, double * vf_diag );

/* Y = alpha * L X + beta * Y with the LowOrder pre/post operator (nrhs must be 3).
 * Replaces: PseudoLaplacian's MultiplyMetric (reference src/Metrics/MetricDimRestricted.hpp:109-139 with
 * src/Metrics/PseudoLaplacian.hpp:66-75). */
int rp_pl_multiply( rp_ctx * ctx, double alpha, const double * X, int64_t ldX, double beta, double * Y, int64_t ldY,
                    int nrhs, int nf_flag, int ff_flag, int ptr_flags );

/* Supernodal Cholesky factor H = L L^T of the H^1 metric, computed on the host: perm (new -> old),
 * supernode column ranges sn_begin (ns+1), sorted row sets rows[row_ptr[s]..row_ptr[s+1]) starting with
 * the supernode's own columns, column-major r x w value blocks at val_ptr[s] whose top w x w block holds
 * the inverse of the diagonal block, and the supernodal elimination-tree parents (postorder, -1 = root).
 * Replaces: the CholeskyDecomposition held by SimplicialMesh::H1Solver (reference
 * src/SimplicialMesh/H1Metric.hpp:208-270). */
int rp_upload_h1_factor( rp_ctx * ctx, int32_t n, int32_t ns, const int32_t * perm, const int32_t * sn_begin,
                         const int64_t * row_ptr, const int32_t * rows, const int64_t * val_ptr, const double * vals,
                         const int32_t * sn_parent );

/* New values (val_ptr[ns] doubles, the layout above) for the factor structure the last rp_upload_h1_factor
 * uploaded: the numeric refactorisation of a moved mesh, whose H^1 metric keeps its pattern, reuses every
 * structure the device built from the symbolic factor (kept until rp_upload_mesh).
 * Replaces: Solver_T::NumericFactorization inside SimplicialMesh::H1Solver (reference
 * src/SimplicialMesh/H1Metric.hpp:256). */
int rp_update_h1_values( rp_ctx * ctx, const double * vals );

/* X = alpha * H^-1 B + beta * X (nrhs must be 3).
 * Replaces: H1Solver().Solve<3>(alpha, B, ldB, beta, X, ldX, nrhs) (reference src/Metrics/TangentPointMetric0.hpp:104-124). */
int rp_h1_solve( rp_ctx * ctx, double alpha, const double * B, int64_t ldB, double beta, double * X, int64_t ldX,
                 int nrhs, int ptr_flags );

/* Y = alpha * H^-1 L_{2-s} H^-1 X + beta * Y (needs rp_upload_h1_factor and rp_pl_compute(2 - s)).
 * Replaces: TangentPointMetric0::multiply_preconditioner (reference src/Metrics/TangentPointMetric0.hpp:88-126). */
int rp_multiply_preconditioner( rp_ctx * ctx, double alpha, const double * X, int64_t ldX, double beta, double * Y, int64_t ldY,
                                int nrhs, int ptr_flags );

/* Preconditioner used by rp_multiply_preconditioner and rp_solve: P = P0 + w diag(A)^-1.
 * w = 0 (default): the reference preconditioner P0 alone. w > 0: fixed weight. w < 0: |w| times an
 * estimate lambda_max(P0 A) / 4 made once per metric/factor state (power iteration, timed as slot 10);
 * w = -1 is the estimate itself.
 * The diagonal term is an addition of this library: P0 under-weights the roughest modes, so the
 * condition number of P0 A grows under refinement; the additive diagonal (Jacobi) term bounds it.
 * No reference counterpart. */
int rp_set_preconditioner( rp_ctx * ctx, double diagonal_weight );

/* Weight w in use (after estimation). */
int rp_preconditioner_weight( rp_ctx * ctx, double * weight );

/* Exact diagonal of the assembled metric A (V values; A acts identically on the three columns).
 * No reference counterpart (used by the diagonal term of the preconditioner). */
int rp_download_metric_diagonal( rp_ctx * ctx, double * diag );

/* Learned preconditioner P = D^{1/2} (P0 + w diag(A)^-1) D^{1/2} (no reference counterpart; the
 * reference preconditioner P0 stays the fallback). A 2-layer MLP of width 64 per cluster-tree node maps
 * the node's features (C x nfeat floats, tree order) to s_c = softplus(o_c) > 0; the vertex scaling is
 * d_v = exp(mean over the primitives containing v of the sum of log s_c along their leaf-to-root path).
 * params: W1 (64 x nfeat, row-major), b1 (64), w2 (64), b2 (1), log w (1): 64 nfeat + 130 floats.
 * The diagonal weight in use is exp(log w) times the base weight: the rp_set_preconditioner weight
 * if positive, else the estimate lambda_max(P0 A) / 4.
 * bf16 != 0 runs inference with BF16 weights and features. nfeat = 0 switches the model off. */
int rp_pc_set_model( rp_ctx * ctx, int nfeat, const float * features, const float * params, int bf16 );

/* Vertex scalings sqrt(d_v) of the current model (V values). */
int rp_pc_download_scaling( rp_ctx * ctx, double * sqrt_d );

/* One training step of the model on the pairs (r_k, e_k = x_last - x_k) recorded by the last
 * rp_solve with RP_SOLVE_RECORD (residuals down to 1e-6 |r_0|): with y_k = r_k - r_last = A e_k,
 * loss = min over alpha of mean_k |alpha P r_k - e_k|_A^2 / (e_k . y_k), in [0, 1] and invariant under
 * scaling P (as PCG is); gradient by HIP
 * kernels (FP32 MLP forward, FP64 reductions), then one Adam step (lr > 0; lr = 0 only computes the
 * gradient) on params / m / v (64 nfeat + 130 floats each, updated in place; step counts from 1).
 * loss and grad (64 nfeat + 130 doubles) may be NULL. */
int rp_pc_learn_step( rp_ctx * ctx, float lr, int step, float * params, float * m, float * v, double * loss, double * grad );

/* rp_solve flags */
#define RP_SOLVE_NO_PRECONDITIONER 1
#define RP_SOLVE_PC_NORM           2
#define RP_SOLVE_GMRES             4
#define RP_SOLVE_RECORD            8   /* PCG: keep r_k and x_k of every iteration for rp_pc_learn_step */

/* X = alpha * A^+ B + beta * X on the three columns of B (zero initial guess).
 * Default: preconditioned conjugate gradients, stopping when every column has
 * |B_k - A Z_k|_2 / |B_k|_2 <= tolerance (recursive residual) or, with RP_SOLVE_PC_NORM, the same ratio
 * in the preconditioner norm |r|_P = sqrt(<r, P r>). With a metric assembled at RP_MIXED_REFINED the
 * iterations multiply with the FP32 copies until that test passes, then the residual is replaced by
 * B - A X with the FP64 metric; the solve ends when the replaced residual passes (the reported measure
 * is then the FP64 one), else it restarts CG from X on the FP64 metric.
 * RP_SOLVE_GMRES: left-preconditioned restarted GMRES(100), stopping when |P (B_k - A Z_k)|_2 / |P B_k|_2
 * <= tolerance (the reference's algorithm and measure).
 * After max_iter iterations without convergence: RP_ESTAGNATION, X still written. iterations and
 * rel_residuals (3 values, in the chosen measure) may be NULL.
 * Replaces: MetricDimRestricted::Solve (reference src/Metrics/MetricDimRestricted.hpp:232-289; the
 * reference runs left-preconditioned GMRES there with the same operator and preconditioner) and
 * Solver_IterationCount / Solver_RelativeResiduals (src/Metrics/MetricBase.hpp:126-134). */
int rp_solve( rp_ctx * ctx, double alpha, const double * B, int64_t ldB, double beta, double * X, int64_t ldX,
              int nrhs, int max_iter, double tolerance, int solver_flags, int ptr_flags, int * iterations, double * rel_residuals );

/* Residual guard of the PCG in rp_solve (no reference counterpart; the learned preconditioner's early
 * fallback uses it): at iteration check_iteration a solve whose largest per-column relative residual (the
 * solve's own measure) is above max_relative_residual stops and returns RP_ESTAGNATION with the iterate
 * reached, so that the caller can switch preconditioner. check_iteration = 0 disables it (the default);
 * the setting holds for later solves on the context until changed. GMRES ignores it. */
int rp_set_solve_guard( rp_ctx * ctx, int check_iteration, double max_relative_residual );

/* Device timing of the last call of each entry point, in milliseconds (HIP events on the
 * context's stream). which: 0 = tp0_compute, 1 = assemble_vertex_gradient, 2 = multiply_metric,
 * 3 = solve, 4 = near kernel, 5 = far kernel, 6 = near SpMM, 7 = multiply_preconditioner,
 * 8 = h1_solve, 9 = pl_compute, 10 = preconditioner setup (diagonal, weight estimate),
 * 11 = very-near field of tp0_compute, 12 = pair stage of max_safe_step, 13 = update_coordinates,
 * 14 = build_topology. */
int rp_last_timing( rp_ctx * ctx, int which, double * ms );

#ifdef __cplusplus
}
#endif

#endif /* REPULSOR_HIP_H */
