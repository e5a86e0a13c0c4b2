// Data preparation, derivative collection, chain rule to vertices, and the cluster-tree passes and
// block SpMMs of MultiplyMetric.
//
// Reference counterparts:
//   k_prep_prims / k_prep_clusters   FMM_Kernel_NF::loadS (barycenter = sum_k (1/(n+1)) hull_k, FMM_Kernel_NF.hpp:169-196),
//                                    FMM_Kernel_FF::loadS (FMM_Kernel_FF.hpp:162-179)
//   k_leaf_far                       CollectFarFieldDerivatives: PercolateDown + ClustersToPrimitives (ClusterTree.hpp:1164-1208)
//   k_hulls                          DNearFarToHulls<2,3>/<1,3> (SimplicialMeshDetails.hpp:529-736, 1097-1410)
//   k_assemble                       DerivativeAssembler().Dot<3> (SimplicialMesh.hpp:625-675)
//   k_pre, k_up_leaves, k_up_level   ClusterTree::Pre: mi_pre.Dot, PrimitivesToClusters, PercolateUp (ClusterTree.hpp:770-857)
//   k_ff_spmm                        TP0_Traversor::FF_MultiplyMetric + DiagonalKernelMatrix (TP0_Traversor.hpp:480-505)
//   k_leaf_path_sum, k_post_*        ClusterTree::Post: PercolateDown, ClustersToPrimitives(add), mi_post.Dot (ClusterTree.hpp:860-943)
// (the near-field SpMM with the fused post epilogue lives in k_nf.hip)

#include "rp_launch.hpp"

namespace rp
{
    // ---------------------------------------------------------------- data preparation
    // Primitive data in structure-of-arrays form (a, x, P) for the row side of the pair kernels, and
    // the same ten values as one record per primitive (rec: a, x, P) for their column gathers.
    __global__ void k_prep_prims( int64_t T, int nd, int size, const double * near, double * a, double * x, double * P, double * rec )
    {
        const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if( i >= T ) { return; }
        const double * X = near + (int64_t)nd * i;
        const double s = 1.0 / size;
        double r[10];
        r[0] = X[0];
        a[i] = r[0];
        for( int k = 0; k < 3; ++k )
        {
            double xk = s * X[1 + k];
            for( int h = 1; h < size; ++h ) { xk += s * X[1 + 3 * h + k]; }
            x[k * T + i] = xk;
            r[1 + k] = xk;
        }
        for( int k = 0; k < 6; ++k ) { P[k * T + i] = X[1 + 3 * size + k]; r[4 + k] = X[1 + 3 * size + k]; }
        for( int k = 0; k < 10; ++k ) { rec[10 * i + k] = r[k]; }
    }

    __global__ void k_prep_clusters( int64_t C, const double * far, double * a, double * x, double * P )
    {
        const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if( i >= C ) { return; }
        const double * X = far + 10 * i;
        a[i] = X[0];
        for( int k = 0; k < 3; ++k ) { x[k * C + i] = X[1 + k]; }
        for( int k = 0; k < 6; ++k ) { P[k * C + i] = X[4 + k]; }
    }

    __global__ void k_leaf_origin( int64_t L, int64_t C, const int32_t * leaf_cluster, const double * cx, double * origin )
    {
        const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if( l >= L ) { return; }
        const int32_t c = leaf_cluster[l];
        for( int k = 0; k < 3; ++k ) { origin[k * L + l] = cx[k * C + c]; }
    }

    // ---------------------------------------------------------------- derivative collection
    // Root-to-leaf sum of the far-field moment derivatives, in the order of the reference's
    // PercolateDown (child = parent + child, parents first). Output 10 x L.
    template<int DIM>
    __global__ void k_leaf_path_sum( int64_t L, int64_t C, const int32_t * anc, const int32_t * anc_len,
                                     const double * src, int64_t src_stride_soa, double * dst, int dst_soa )
    {
        const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if( l >= L ) { return; }
        // the leaf's path from the ancestor table built at tree upload, anc[d * L + l] = the cluster at
        // depth d (d = 0: the root): every index load is independent of the sums, so they go out ahead
        const int n = anc_len[l];
        // SoA source: src[k * C + c]; AoS source: src[DIM * c + k]
        auto val = [&]( int32_t c, int k ) -> double { return src_stride_soa ? src[k * C + c] : src[(int64_t)DIM * c + k]; };
        double acc[DIM];
        const int32_t r = anc[l];
        for( int k = 0; k < DIM; ++k ) { acc[k] = val( r, k ); }
        #pragma unroll 1
        for( int d = 1; d < n; ++d )
        {
            const int32_t cc = anc[(int64_t)d * L + l];
            for( int k = 0; k < DIM; ++k ) { acc[k] = acc[k] + val( cc, k ); }
        }
        for( int k = 0; k < DIM; ++k )
        {
            if( dst_soa ) { dst[k * L + l] = acc[k]; } else { dst[(int64_t)DIM * l + k] = acc[k]; }
        }
    }

    // Chain rule from the moment derivatives (near: a, a*hull, a*P; far: a, a*x, a*P) to the vertex
    // coordinates of every simplex. With w = normal (X1-X0)x(X2-X0) for triangles or tangent X1-X0
    // for edges, vol = |w|/2 or |w|, and M the symmetric matrix of the projector weights
    // W_k = N_P,k + F_P,k (M_aa = W_aa, M_ab = W_ab/2):
    //   triangles: pr = w w^T/|w|^2, G = d(vol*K + vol*W.pr)/dw = ((K - S) w + 2 M w) / (2|w|), S = w.Mw/|w|^2
    //   edges:     pr = I - w w^T/|w|^2, G = ((K + tr W - w.Mw/|w|^2 + 2 w.Mw/|w|^2) w - 2 M w) / |w|
    //   K = N_0 + F_0 + sum_l N_hull,l . X_l + F_x . mean(X)
    //   dB/dX_{j,c} = G . dw/dX_{j,c} + vol (N_hull,j,c + F_x,c / (n+1))
    // with dw/dX_{j,c} = e_c x u_j (u_0 = e1 - e2, u_1 = e2, u_2 = -e1) for triangles and -/+ e_c for edges.
    template<int N>
    __global__ void k_hulls( int64_t T, int64_t L, const int32_t * P_ordering, const int32_t * leaf_of_prim,
                             const int32_t * S, const double * V, const double * charges,
                             const double * nf_dx, const double * vf_dh, const double * leaf_far, double * B )
    {
        constexpr int SIZE = N + 1;
        const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if( i >= T ) { return; }
        const int64_t t = P_ordering[i];
        const int64_t l = leaf_of_prim[i];
        double X[SIZE][3];
        double charge = 0;
        for( int j = 0; j < SIZE; ++j )
        {
            const int32_t v = S[SIZE * t + j];
            for( int c = 0; c < 3; ++c ) { X[j][c] = V[3 * (int64_t)v + c]; }
            charge += charges ? charges[v] : 1.0;
        }
        charge *= 1.0 / SIZE;
        double Nd[10], Fd[10];
        for( int k = 0; k < 10; ++k ) { Nd[k] = nf_dx[k * T + i]; Fd[k] = leaf_far[k * L + l]; }
        // Nd: [0] d/da, [1..3] d/d(a hull_l) (same for every l), [4..9] d/d(a P)
        double mean[3] = { 0, 0, 0 };
        for( int j = 0; j < SIZE; ++j ) { for( int c = 0; c < 3; ++c ) { mean[c] += X[j][c] * (1.0 / SIZE); } }
        // very-near-field hull-vertex terms (they differ per vertex), zero without a very-near field
        double Hv[SIZE][3];
        for( int j = 0; j < SIZE; ++j ) { for( int c = 0; c < 3; ++c ) { Hv[j][c] = vf_dh ? vf_dh[3 * SIZE * i + 3 * j + c] : 0.0; } }
        double K = Nd[0] + Fd[0];
        for( int j = 0; j < SIZE; ++j ) { for( int c = 0; c < 3; ++c ) { K += ( Nd[1 + c] + Hv[j][c] ) * X[j][c]; } }
        for( int c = 0; c < 3; ++c ) { K += Fd[1 + c] * mean[c]; }
        double W[6];
        for( int k = 0; k < 6; ++k ) { W[k] = Nd[4 + k] + Fd[4 + k]; }
        double w[3];
        const double e1[3] = { X[1][0] - X[0][0], X[1][1] - X[0][1], X[1][2] - X[0][2] };
        double e2[3] = { 0, 0, 0 };
        if constexpr ( N == 2 )
        {
            for( int c = 0; c < 3; ++c ) { e2[c] = X[2][c] - X[0][c]; }
            w[0] = e1[1] * e2[2] - e1[2] * e2[1];
            w[1] = e1[2] * e2[0] - e1[0] * e2[2];
            w[2] = e1[0] * e2[1] - e1[1] * e2[0];
        }
        else
        {
            for( int c = 0; c < 3; ++c ) { w[c] = e1[c]; }
        }
        const double ww = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
        const double nn = sqrt( ww );
        const double Mw[3] = {
            W[0] * w[0] + 0.5 * W[1] * w[1] + 0.5 * W[2] * w[2],
            0.5 * W[1] * w[0] + W[3] * w[1] + 0.5 * W[4] * w[2],
            0.5 * W[2] * w[0] + 0.5 * W[4] * w[1] + W[5] * w[2] };
        const double wMw = w[0] * Mw[0] + w[1] * Mw[1] + w[2] * Mw[2];
        double G[3], vol;
        if constexpr ( N == 2 )
        {
            vol = 0.5 * nn;
            const double Sv = wMw / ww;
            for( int c = 0; c < 3; ++c ) { G[c] = ( (K - Sv) * w[c] + 2.0 * Mw[c] ) / (2.0 * nn); }
        }
        else
        {
            vol = nn;
            const double trW = W[0] + W[3] + W[5];
            const double Qv = wMw / ww;
            const double Sedge = trW - Qv;
            for( int c = 0; c < 3; ++c ) { G[c] = ( (K + Sedge + 2.0 * Qv) * w[c] - 2.0 * Mw[c] ) / nn; }
        }
        for( int j = 0; j < SIZE; ++j )
        {
            double d[3];
            if constexpr ( N == 2 )
            {
                double u[3];
                for( int c = 0; c < 3; ++c ) { u[c] = (j == 0) ? (e1[c] - e2[c]) : ((j == 1) ? e2[c] : -e1[c]); }
                // (u x G)_c = G . (e_c x u)
                d[0] = u[1] * G[2] - u[2] * G[1];
                d[1] = u[2] * G[0] - u[0] * G[2];
                d[2] = u[0] * G[1] - u[1] * G[0];
            }
            else
            {
                const double sgn = (j == 0) ? -1.0 : 1.0;
                for( int c = 0; c < 3; ++c ) { d[c] = sgn * G[c]; }
            }
            for( int c = 0; c < 3; ++c )
            {
                B[ (SIZE * t + j) * 3 + c ] = charge * ( d[c] + vol * ( Nd[1 + c] + Hv[j][c] + Fd[1 + c] * (1.0 / SIZE) ) );
            }
        }
    }

    // Y[v] = alpha * sum_{corners r of v, increasing} B[r] + beta * Y[v]
    __global__ void k_assemble( int64_t V, const int64_t * vc_ptr, const int32_t * vc_corner, const double * B,
                                double alpha, double beta, double * Y, int64_t ldY )
    {
        const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if( v >= V ) { return; }
        double s0 = 0, s1 = 0, s2 = 0;
        for( int64_t k = vc_ptr[v]; k < vc_ptr[v + 1]; ++k )
        {
            const int64_t r = vc_corner[k];
            s0 += B[3 * r]; s1 += B[3 * r + 1]; s2 += B[3 * r + 2];
        }
        double * y = Y + ldY * v;
        if( beta == 0.0 ) { y[0] = alpha * s0; y[1] = alpha * s1; y[2] = alpha * s2; }
        else { y[0] = alpha * s0 + beta * y[0]; y[1] = alpha * s1 + beta * y[1]; y[2] = alpha * s2 + beta * y[2]; }
    }

    // ---------------------------------------------------------------- MultiplyMetric: Pre
    // P_in[i][r][k] = sum_l pre_vals[i][r][l] X[pre_cols[i][l]][k]   (4 x 3 per primitive)
    template<int SIZE>
    __global__ void k_pre( int64_t T, const int32_t * cols, const double * vals, const double * X, int64_t ldX, double * P_in )
    {
        const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if( i >= T ) { return; }
        double x[SIZE][3];
        for( int l = 0; l < SIZE; ++l )
        {
            const int64_t v = cols[SIZE * i + l];
            for( int k = 0; k < 3; ++k ) { x[l][k] = X[ldX * v + k]; }
        }
        const double * pv = vals + 4 * SIZE * i;
        double out[12];
        for( int r = 0; r < 4; ++r )
        {
            for( int k = 0; k < 3; ++k )
            {
                double s = 0;
                for( int l = 0; l < SIZE; ++l ) { s += pv[SIZE * r + l] * x[l][k]; }
                out[3 * r + k] = s;
            }
        }
        double2 * o = reinterpret_cast<double2 *>( P_in + 12 * i );
        for( int h = 0; h < 6; ++h ) { o[h] = make_double2( out[2 * h], out[2 * h + 1] ); }
    }

    // Leaf sums of P_in (PrimitivesToClusters), W values per primitive.
    __global__ void k_up_leaves( int64_t L, int W, const int32_t * leaf_cluster, const int32_t * C_begin, const int32_t * C_end,
                                 const double * P_in, double * C_in )
    {
        const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        const int64_t l = g / W;
        const int h = (int)(g % W);
        if( l >= L ) { return; }
        const int32_t c = leaf_cluster[l];
        // the leaf's values in batches of 8 loads (leaves hold at most 32 primitives), summed in order
        const int32_t b = C_begin[c], e = C_end[c];
        double s = 0;
        for( int32_t i0 = b; i0 < e; i0 += 8 )
        {
            double v[8];
            #pragma unroll
            for( int u = 0; u < 8; ++u ) { v[u] = ( i0 + u < e ) ? P_in[W * (int64_t)( i0 + u ) + h] : 0.0; }
            #pragma unroll
            for( int u = 0; u < 8; ++u ) { if( i0 + u < e ) { s += v[u]; } }
        }
        C_in[W * (int64_t)c + h] = s;
    }

    // One depth level of PercolateUp: C_in[c] = C_in[left] + C_in[right].
    __global__ void k_up_level( int64_t n, int W, const int32_t * nodes, const int32_t * C_left, const int32_t * C_right, double * C_in )
    {
        const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        const int64_t q = g / W;
        const int h = (int)(g % W);
        if( q >= n ) { return; }
        const int32_t c = nodes[q];
        C_in[W * (int64_t)c + h] = C_in[W * (int64_t)C_left[c] + h] + C_in[W * (int64_t)C_right[c] + h];
    }

    // Three depth levels of PercolateUp in one launch: one thread per (internal node c at the band's top
    // depth, component h) computes C_in = C_in[left] + C_in[right] for c's internal descendants two and
    // one levels down (band size 3, fewer at the root end), then for c: the same sums as one level per
    // launch. Children below the band were done by the previous (deeper) band.
    __global__ void k_up_band( int64_t n, int W, int band, const int32_t * nodes, const int32_t * C_left, const int32_t * C_right, double * C_in )
    {
        const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        const int64_t q = g / W;
        const int h = (int)(g % W);
        if( q >= n ) { return; }
        const int32_t c = nodes[q];
        int32_t l1[2], l2[4];
        int n1 = 0, n2 = 0;
        if( band >= 2 )
        {
            const int32_t a = C_left[c], b = C_right[c];
            if( C_left[a] >= 0 ) { l1[n1++] = a; }
            if( C_left[b] >= 0 ) { l1[n1++] = b; }
            if( band >= 3 )
            {
                for( int u = 0; u < n1; ++u )
                {
                    const int32_t x = C_left[l1[u]], y = C_right[l1[u]];
                    if( C_left[x] >= 0 ) { l2[n2++] = x; }
                    if( C_left[y] >= 0 ) { l2[n2++] = y; }
                }
            }
        }
        for( int u = 0; u < n2; ++u )
        {
            const int32_t x = l2[u];
            C_in[W * (int64_t)x + h] = C_in[W * (int64_t)C_left[x] + h] + C_in[W * (int64_t)C_right[x] + h];
        }
        for( int u = 0; u < n1; ++u )
        {
            const int32_t x = l1[u];
            C_in[W * (int64_t)x + h] = C_in[W * (int64_t)C_left[x] + h] + C_in[W * (int64_t)C_right[x] + h];
        }
        C_in[W * (int64_t)c + h] = C_in[W * (int64_t)C_left[c] + h] + C_in[W * (int64_t)C_right[c] + h];
    }

    // The top levels of PercolateUp in one workgroup: every level whose internal clusters number at most
    // UP_TOP_MAX, deepest first, one barrier between levels (the sums of a level are written by the
    // workgroup's own threads and read after the barrier), the same pair sums as one level per launch.
    constexpr int UP_TOP_MAX = 512;
    constexpr int UP_TOP_LEVELS = 48;
    struct UpTopLevels { int n; int32_t ptr[UP_TOP_LEVELS + 1]; };

    __global__ __launch_bounds__(1024) void k_up_top( UpTopLevels lv, int W, const int32_t * nodes, const int32_t * C_left, const int32_t * C_right, double * C_in )
    {
        for( int l = 0; l < lv.n; ++l )
        {
            const int32_t b = lv.ptr[l], cnt = ( lv.ptr[l + 1] - b ) * W;
            for( int32_t g = threadIdx.x; g < cnt; g += 1024 )
            {
                const int32_t c = nodes[b + g / W];
                const int h = g % W;
                C_in[W * (int64_t)c + h] = C_in[W * (int64_t)C_left[c] + h] + C_in[W * (int64_t)C_right[c] + h];
            }
            __syncthreads();
        }
    }

    template<typename T>
    __device__ __forceinline__ T wave_sum64( T v )
    {
        #pragma unroll
        for( int off = 32; off > 0; off >>= 1 ) { v += __shfl_xor( v, off, 64 ); }
        return v;
    }

    // Far-field block SpMM: a group of FF_G lanes per cluster row (rows hold about two dozen entries),
    // group sums by lane shuffles, then the diagonal block.
    constexpr int FF_G = 16;

    template<typename VT>
    __global__ __launch_bounds__(256) void k_ff_spmm( int64_t C, const int64_t * rp, const int32_t * ci,
        const VT * vals, const double * diag, const double * xin, double * yout )
    {
        // workgroups dealt to the XCDs round robin; each XCD takes one contiguous run of cluster rows
        // (subtrees of the cluster tree, whose far partners overlap), so the gathered input blocks are
        // reused in that XCD's L2
        const int64_t per = gridDim.x / NUM_XCD;
        const int64_t blk = (int64_t)( blockIdx.x % NUM_XCD ) * per + blockIdx.x / NUM_XCD;
        const int lane = threadIdx.x % FF_G;
        const int64_t i = ( blk * 256 + threadIdx.x ) / FF_G;
        if( i >= C ) { return; }
        double acc[12];
        #pragma unroll
        for( int k = 0; k < 12; ++k ) { acc[k] = 0.0; }
        // a lane takes the entries k, k + FF_G, ... of the row in that order, two per step with both
        // entries' indices, values and input blocks loaded before either product (rows hold about two
        // dozen entries, so most lanes need one step: one index -> block round trip instead of two)
        const int64_t e = rp[i + 1];
        auto fma_block = [&]( const VT av, const double * x )
        {
            const double a0 = (double)av.x, a1 = (double)av.y;
            #pragma unroll
            for( int h = 0; h < 3; ++h ) { acc[h] += a0 * x[h]; }
            #pragma unroll
            for( int h = 3; h < 12; ++h ) { acc[h] += a1 * x[h]; }
        };
        for( int64_t k = rp[i] + lane; k < e; k += 2 * FF_G )
        {
            const bool two = k + FF_G < e;
            const int32_t j0 = ci[k];
            const VT av0 = vals[k];
            int32_t j1 = j0;
            VT av1 = av0;
            if( two ) { j1 = ci[k + FF_G]; av1 = vals[k + FF_G]; }
            const double2 * x0p = reinterpret_cast<const double2 *>( xin + 12 * (int64_t)j0 );
            const double2 * x1p = reinterpret_cast<const double2 *>( xin + 12 * (int64_t)j1 );
            double x0[12], x1[12];
            #pragma unroll
            for( int h = 0; h < 6; ++h ) { const double2 t = x0p[h]; x0[2 * h] = t.x; x0[2 * h + 1] = t.y; }
            #pragma unroll
            for( int h = 0; h < 6; ++h ) { const double2 t = x1p[h]; x1[2 * h] = t.x; x1[2 * h + 1] = t.y; }
            fma_block( av0, x0 );
            if( two ) { fma_block( av1, x1 ); }
        }
        #pragma unroll
        for( int k = 0; k < 12; ++k )
        {
            #pragma unroll
            for( int off = FF_G / 2; off > 0; off >>= 1 ) { acc[k] += __shfl_xor( acc[k], off, 64 ); }
        }
        if( lane == 0 )
        {
            const double d0 = diag[2 * i], d1 = diag[2 * i + 1];
            const double * xi = xin + 12 * i;
            for( int k = 0; k < 3; ++k )  { acc[k] += d0 * xi[k]; }
            for( int k = 3; k < 12; ++k ) { acc[k] += d1 * xi[k]; }
            for( int k = 0; k < 12; ++k ) { yout[12 * i + k] = acc[k]; }
        }
    }

    // ---------------------------------------------------------------- MultiplyMetric: Post
    // Primitive side of the post-operator when the near-field SpMM is skipped: ytot_i = far-field
    // leaf total of leaf(i) (or 0), S_out[i][s][c] = sum_r pre_vals[i][r][s] ytot_i[r][c].
    template<int SIZE>
    __global__ void k_post_prim( int64_t T, const int32_t * leaf_idx_of_prim, const double * pre_vals,
                                 const double * leaf_out, int use_far, double * S_out )
    {
        const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if( i >= T ) { return; }
        double s[12];
        for( int k = 0; k < 12; ++k ) { s[k] = use_far ? leaf_out[12 * (int64_t)leaf_idx_of_prim[i] + k] : 0.0; }
        const double * pv = pre_vals + 4 * SIZE * i;
        for( int sl = 0; sl < SIZE; ++sl )
        {
            for( int cc = 0; cc < 3; ++cc )
            {
                double t = 0.0;
                for( int r = 0; r < 4; ++r ) { t += pv[SIZE * r + sl] * s[3 * r + cc]; }
                S_out[(SIZE * i + sl) * 3 + cc] = t;
            }
        }
    }

    // Y[v] = alpha * sum_{(i,slot) of v, increasing i} S_out[i][slot] + beta * Y[v]   (mi_post.Dot)
    // perm non-null: row j of Y takes vertex perm[j] (the H^1 factor's numbering inside the preconditioner)
    __global__ void k_post_vert( int64_t V, const int64_t * vp_ptr, const int32_t * vp_entry, const double * S_out,
                                 double alpha, double beta, double * Y, int64_t ldY, const int32_t * perm )
    {
        const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if( j >= V ) { return; }
        const int64_t v = perm ? (int64_t)perm[j] : j;
        double s0 = 0, s1 = 0, s2 = 0;
        const int64_t kb = vp_ptr[v], ke = vp_ptr[v + 1];
        // up to eight entries (a vertex's valence is about six) are gathered at once, then summed in order
        for( int64_t k0 = kb; k0 < ke; k0 += 8 )
        {
            int32_t e[8];
            #pragma unroll
            for( int u = 0; u < 8; ++u ) { e[u] = ( k0 + u < ke ) ? vp_entry[k0 + u] : -1; }
            double t[8][3];
            #pragma unroll
            for( int u = 0; u < 8; ++u )
            {
                if( e[u] >= 0 ) { const double * so = S_out + 3 * (int64_t)e[u]; t[u][0] = so[0]; t[u][1] = so[1]; t[u][2] = so[2]; }
                else { t[u][0] = t[u][1] = t[u][2] = 0.0; }
            }
            #pragma unroll
            for( int u = 0; u < 8; ++u ) { if( e[u] >= 0 ) { s0 += t[u][0]; s1 += t[u][1]; s2 += t[u][2]; } }
        }
        double * y = Y + ldY * j;
        if( beta == 0.0 ) { y[0] = alpha * s0; y[1] = alpha * s1; y[2] = alpha * s2; }
        else { y[0] = alpha * s0 + beta * y[0]; y[1] = alpha * s1 + beta * y[1]; y[2] = alpha * s2 + beta * y[2]; }
    }

    __global__ void k_map_index( int64_t n, const int32_t * map, const int32_t * in, int32_t * out )
    {
        const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if( g < n ) { out[g] = map[in[g]]; }
    }

    __global__ void k_fill( double * p, int64_t n, double v )
    {
        const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if( g < n ) { p[g] = v; }
    }

    // ---------------------------------------------------------------- host launchers
    static inline dim3 grid1( int64_t n, int b = 256 ) { return dim3( (unsigned)std::max<int64_t>( 1, (n + b - 1) / b ) ); }

    void launch_prep_prims( int64_t T, int nd, int size, const double * near, double * a, double * x, double * P, double * rec, hipStream_t s )
    { if( T ) { hipLaunchKernelGGL( k_prep_prims, grid1(T), dim3(256), 0, s, T, nd, size, near, a, x, P, rec ); } }

    void launch_prep_clusters( int64_t C, const double * far, double * a, double * x, double * P, hipStream_t s )
    { if( C ) { hipLaunchKernelGGL( k_prep_clusters, grid1(C), dim3(256), 0, s, C, far, a, x, P ); } }

    void launch_leaf_origin( int64_t L, int64_t C, const int32_t * leaf_cluster, const double * cx, double * origin, hipStream_t s )
    { if( L ) { hipLaunchKernelGGL( k_leaf_origin, grid1(L), dim3(256), 0, s, L, C, leaf_cluster, cx, origin ); } }

    void launch_leaf_far( int64_t L, int64_t C, const int32_t * leaf_anc, const int32_t * leaf_anc_len, const double * ff_dx, double * leaf_far, hipStream_t s )
    { if( L ) { hipLaunchKernelGGL( (k_leaf_path_sum<10>), grid1(L), dim3(256), 0, s, L, C, leaf_anc, leaf_anc_len, ff_dx, (int64_t)1, leaf_far, 1 ); } }

    void launch_leaf_out( int64_t L, int64_t C, const int32_t * leaf_anc, const int32_t * leaf_anc_len, const double * C_out, double * leaf_out, hipStream_t s )
    { if( L ) { hipLaunchKernelGGL( (k_leaf_path_sum<12>), grid1(L), dim3(256), 0, s, L, C, leaf_anc, leaf_anc_len, C_out, (int64_t)0, leaf_out, 0 ); } }

    void launch_hulls( int n, int64_t T, int64_t L, const int32_t * P_ordering, const int32_t * leaf_of_prim, const int32_t * S,
                       const double * V, const double * charges, const double * nf_dx, const double * vf_dh, const double * leaf_far, double * B, hipStream_t s )
    {
        if( !T ) { return; }
        if( n == 2 ) { hipLaunchKernelGGL( (k_hulls<2>), grid1(T), dim3(256), 0, s, T, L, P_ordering, leaf_of_prim, S, V, charges, nf_dx, vf_dh, leaf_far, B ); }
        else         { hipLaunchKernelGGL( (k_hulls<1>), grid1(T), dim3(256), 0, s, T, L, P_ordering, leaf_of_prim, S, V, charges, nf_dx, vf_dh, leaf_far, B ); }
    }

    void launch_assemble( int64_t V, const int64_t * vc_ptr, const int32_t * vc_corner, const double * B, double alpha, double beta, double * Y, int64_t ldY, hipStream_t s )
    { if( V ) { hipLaunchKernelGGL( k_assemble, grid1(V), dim3(256), 0, s, V, vc_ptr, vc_corner, B, alpha, beta, Y, ldY ); } }

    void launch_pre( int size, int64_t T, const int32_t * cols, const double * vals, const double * X, int64_t ldX, double * P_in, hipStream_t s )
    {
        if( !T ) { return; }
        if( size == 3 ) { hipLaunchKernelGGL( (k_pre<3>), grid1(T), dim3(256), 0, s, T, cols, vals, X, ldX, P_in ); }
        else            { hipLaunchKernelGGL( (k_pre<2>), grid1(T), dim3(256), 0, s, T, cols, vals, X, ldX, P_in ); }
    }

    void launch_up_leaves( int64_t L, const int32_t * leaf_cluster, const int32_t * C_begin, const int32_t * C_end, const double * P_in, double * C_in, hipStream_t s, int W )
    { if( L ) { hipLaunchKernelGGL( k_up_leaves, grid1(W * L), dim3(256), 0, s, L, W, leaf_cluster, C_begin, C_end, P_in, C_in ); } }

    void launch_up_level( int64_t n, const int32_t * nodes, const int32_t * C_left, const int32_t * C_right, double * C_in, hipStream_t s, int W )
    { if( n ) { hipLaunchKernelGGL( k_up_level, grid1(W * n), dim3(256), 0, s, n, W, nodes, C_left, C_right, C_in ); } }

    void launch_up_tree( const std::vector<int32_t> & level_ptr, const int32_t * level_nodes, const int32_t * C_left, const int32_t * C_right,
                         double * C_in, hipStream_t s, int W )
    {
        // level index l holds the internal clusters of depth maxd - l; the levels from top on (at most
        // UP_TOP_MAX clusters each, the root end) run in one workgroup, the deeper ones in bands of three
        // levels ending at l = 2, 5, ... and at top - 1
        const int nl = (int)level_ptr.size() - 1;
        int top = nl;
        while( top > 0 && nl - top < UP_TOP_LEVELS && level_ptr[(size_t)top] - level_ptr[(size_t)top - 1] <= UP_TOP_MAX ) { --top; }
        int prev = -1;
        for( int l = 0; l < top; ++l )
        {
            if( l % 3 != 2 && l != top - 1 ) { continue; }
            const int64_t b = level_ptr[(size_t)l], e = level_ptr[(size_t)l + 1];
            const int band = l - prev;
            prev = l;
            if( e > b ) { hipLaunchKernelGGL( k_up_band, grid1(W * ( e - b )), dim3(256), 0, s, e - b, W, band, level_nodes + b, C_left, C_right, C_in ); }
        }
        if( top < nl )
        {
            UpTopLevels lv;
            lv.n = nl - top;
            for( int l = 0; l <= lv.n; ++l ) { lv.ptr[l] = level_ptr[(size_t)( top + l )]; }
            hipLaunchKernelGGL( k_up_top, dim3(1), dim3(1024), 0, s, lv, W, level_nodes, C_left, C_right, C_in );
        }
    }

    void launch_leaf_out3( int64_t L, int64_t C, const int32_t * leaf_anc, const int32_t * leaf_anc_len, const double * C_out, double * leaf_out, hipStream_t s )
    { if( L ) { hipLaunchKernelGGL( (k_leaf_path_sum<3>), grid1(L), dim3(256), 0, s, L, C, leaf_anc, leaf_anc_len, C_out, (int64_t)0, leaf_out, 0 ); } }

    void launch_leaf_path_sum1( int64_t L, int64_t C, const int32_t * leaf_anc, const int32_t * leaf_anc_len, const double * src, double * leaf_out, hipStream_t s )
    { if( L ) { hipLaunchKernelGGL( (k_leaf_path_sum<1>), grid1(L), dim3(256), 0, s, L, C, leaf_anc, leaf_anc_len, src, (int64_t)0, leaf_out, 0 ); } }

    void launch_ff_spmm( bool mixed, int64_t C, const int64_t * rp, const int32_t * ci, const void * vals, const double * diag, const double * xin, double * yout, hipStream_t s )
    {
        if( !C ) { return; }
        const int64_t blocks = (C + 256 / FF_G - 1) / (256 / FF_G);
        const dim3 grid( (unsigned)( ( ( blocks + NUM_XCD - 1 ) / NUM_XCD ) * NUM_XCD ) );
        if( mixed ) { hipLaunchKernelGGL( (k_ff_spmm<float2>), grid, dim3(256), 0, s, C, rp, ci, (const float2 *)vals, diag, xin, yout ); }
        else        { hipLaunchKernelGGL( (k_ff_spmm<double2>), grid, dim3(256), 0, s, C, rp, ci, (const double2 *)vals, diag, xin, yout ); }
    }

    void launch_post_prim( int size, int64_t T, const int32_t * leaf_idx_of_prim, const double * pre_vals, const double * leaf_out, int use_far, double * S_out, hipStream_t s )
    {
        if( !T ) { return; }
        if( size == 3 ) { hipLaunchKernelGGL( (k_post_prim<3>), grid1(T), dim3(256), 0, s, T, leaf_idx_of_prim, pre_vals, leaf_out, use_far, S_out ); }
        else            { hipLaunchKernelGGL( (k_post_prim<2>), grid1(T), dim3(256), 0, s, T, leaf_idx_of_prim, pre_vals, leaf_out, use_far, S_out ); }
    }

    void launch_post_vert( int64_t V, const int64_t * vp_ptr, const int32_t * vp_entry, const double * S_out, double alpha, double beta, double * Y, int64_t ldY, hipStream_t s,
                           const int32_t * perm )
    { if( V ) { hipLaunchKernelGGL( k_post_vert, grid1(V), dim3(256), 0, s, V, vp_ptr, vp_entry, S_out, alpha, beta, Y, ldY, perm ); } }

    void launch_map_index( int64_t n, const int32_t * map, const int32_t * in, int32_t * out, hipStream_t s )
    { if( n ) { hipLaunchKernelGGL( k_map_index, grid1(n), dim3(256), 0, s, n, map, in, out ); } }

    void launch_fill( double * p, int64_t n, double v, hipStream_t s )
    { if( n ) { hipLaunchKernelGGL( k_fill, grid1(n), dim3(256), 0, s, p, n, v ); } }

} // namespace rp
