// Triangular solves with the supernodal Cholesky factor of the H^1 metric, H = L L^T, for three
// right-hand sides (row-major n x 3). The reference solves with its sparse Cholesky factor on the host
// (src/SimplicialMesh/H1Metric.hpp:208-270, called twice per preconditioner application from
// src/Metrics/TangentPointMetric0.hpp:88-126).
//
// Factor layout (one column-major r x w block per supernode s, rows R_s sorted, the first w of them
// the supernode's own columns f..f+w-1): the top w x w block holds inv(L_ss) (lower triangle), the
// rows below hold L_off = L(R_s[w:], f:f+w).
//
// Forward (multifrontal, leaves to root, one launch per height level): the front vector z_s (length r)
// starts as [b_s; 0], receives the update vectors of the children in child order through the index
// map cmap (position of each child update row in the parent's row list), then
//     y_s = inv(L_ss) z_s[0:w],   z_s[w:r] -= L_off y_s.
// z_s[0:w] keeps y_s for the backward sweep. The scatter of a child's update is conflict-free (its
// rows are distinct) and children are merged one after another, so the sums are deterministic.
// Backward (root to leaves):  x_s = inv(L_ss)^T ( y_s - L_off^T x(R_s[w:]) ).

#include "rp_launch.hpp"

namespace rp
{
    constexpr int CHOL_QC = 256;   // ancestor rows staged in LDS per chunk of the backward sweep

    __device__ __forceinline__ int pow2_at_least( int x )
    {
        int g = 1;
        while( g < x && g < 64 ) { g <<= 1; }
        return g;
    }

    template<int BS>
    __global__ __launch_bounds__(BS) void k_chol_fwd( CholDev F, const int32_t * nodes, double alpha, const double * B, int64_t ldB )
    {
        extern __shared__ double sh[];
        const int s = nodes[blockIdx.x];
        const int f = F.sn_begin[s];
        const int w = F.sn_begin[s + 1] - f;
        const int64_t rb = F.row_ptr[s];
        const int r = (int)( F.row_ptr[s + 1] - rb );
        double * z = F.z + 3 * rb;
        const double * Ls = F.vals + F.val_ptr[s];
        double * zs = sh;            // 3 w
        double * ys = sh + 3 * w;    // 3 w
        const int tid = threadIdx.x;
        for( int a = tid; a < w; a += BS )
        {
            const int64_t o = F.perm[f + a];
            for( int k = 0; k < 3; ++k ) { zs[3 * a + k] = alpha * B[ldB * o + k]; }
        }
        for( int q = w + tid; q < r; q += BS ) { for( int k = 0; k < 3; ++k ) { z[3 * q + k] = 0.0; } }
        __syncthreads();
        for( int kk = F.kid_ptr[s]; kk < F.kid_ptr[s + 1]; ++kk )
        {
            const int c = F.kids[kk];
            const int wc = F.sn_begin[c + 1] - F.sn_begin[c];
            const int64_t rbc = F.row_ptr[c];
            const int rc = (int)( F.row_ptr[c + 1] - rbc );
            const double * zc = F.z + 3 * rbc;
            for( int q = wc + tid; q < rc; q += BS )
            {
                const int pos = F.cmap[rbc + q];
                if( pos < w ) { for( int k = 0; k < 3; ++k ) { zs[3 * pos + k] += zc[3 * q + k]; } }
                else          { for( int k = 0; k < 3; ++k ) { z[3 * pos + k]  += zc[3 * q + k]; } }
            }
            __syncthreads();
        }
        for( int a = tid; a < w; a += BS )
        {
            double acc[3] = { 0.0, 0.0, 0.0 };
            #pragma unroll 8
            for( int c = 0; c <= a; ++c )
            {
                const double l = Ls[(int64_t)c * r + a];
                for( int k = 0; k < 3; ++k ) { acc[k] += l * zs[3 * c + k]; }
            }
            for( int k = 0; k < 3; ++k ) { ys[3 * a + k] = acc[k]; z[3 * a + k] = acc[k]; }
        }
        __syncthreads();
        for( int q = w + tid; q < r; q += BS )
        {
            double acc[3] = { z[3 * q], z[3 * q + 1], z[3 * q + 2] };
            #pragma unroll 8
            for( int c = 0; c < w; ++c )
            {
                const double l = Ls[(int64_t)c * r + q];
                for( int k = 0; k < 3; ++k ) { acc[k] -= l * ys[3 * c + k]; }
            }
            for( int k = 0; k < 3; ++k ) { z[3 * q + k] = acc[k]; }
        }
    }

    // Sum of v over aligned groups of G lanes (G a power of two <= 64); every lane takes part.
    __device__ __forceinline__ double group_sum( double v, int G )
    {
        for( int off = G >> 1; off > 0; off >>= 1 ) { v += __shfl_xor( v, off, 64 ); }
        return v;
    }

    template<int BS>
    __global__ __launch_bounds__(BS) void k_chol_bwd( CholDev F, const int32_t * nodes )
    {
        extern __shared__ double sh[];
        const int s = nodes[blockIdx.x];
        const int f = F.sn_begin[s];
        const int w = F.sn_begin[s + 1] - f;
        const int64_t rb = F.row_ptr[s];
        const int r = (int)( F.row_ptr[s + 1] - rb );
        const double * z = F.z + 3 * rb;
        const int32_t * R = F.rows + rb;
        const double * Ls = F.vals + F.val_ptr[s];
        double * t  = sh;               // 3 w
        double * xr = sh + 3 * w;       // 3 CHOL_QC
        const int tid = threadIdx.x;
        for( int a = tid; a < w; a += BS ) { for( int k = 0; k < 3; ++k ) { t[3 * a + k] = z[3 * a + k]; } }
        const int u = r - w;
        // phase 1: t -= L_off^T x(R[w:]), groups of G1 lanes per column
        if( u > 0 )
        {
            const int G1 = pow2_at_least( u < CHOL_QC ? u : CHOL_QC );
            const int ng = BS / G1;
            const int grp = tid / G1, gl = tid % G1;
            for( int q0 = 0; q0 < u; q0 += CHOL_QC )
            {
                const int qn = (u - q0 < CHOL_QC) ? (u - q0) : CHOL_QC;
                __syncthreads();
                for( int q = tid; q < qn; q += BS )
                {
                    const int64_t j = R[w + q0 + q];
                    for( int k = 0; k < 3; ++k ) { xr[3 * q + k] = F.xp[3 * j + k]; }
                }
                __syncthreads();
                for( int c0 = 0; c0 < w; c0 += ng )
                {
                    const int c = c0 + grp;
                    double acc[3] = { 0.0, 0.0, 0.0 };
                    if( c < w )
                    {
                        const double * col = Ls + (int64_t)c * r + w + q0;
                        #pragma unroll 8
                        for( int q = gl; q < qn; q += G1 )
                        {
                            const double l = col[q];
                            for( int k = 0; k < 3; ++k ) { acc[k] += l * xr[3 * q + k]; }
                        }
                    }
                    for( int k = 0; k < 3; ++k ) { acc[k] = group_sum( acc[k], G1 ); }
                    if( c < w && gl == 0 ) { for( int k = 0; k < 3; ++k ) { t[3 * c + k] -= acc[k]; } }
                }
            }
        }
        __syncthreads();
        // phase 2: x = inv(L_ss)^T t, groups of G2 lanes per column
        {
            const int G2 = pow2_at_least( w );
            const int ng = BS / G2;
            const int grp = tid / G2, gl = tid % G2;
            for( int c0 = 0; c0 < w; c0 += ng )
            {
                const int c = c0 + grp;
                double acc[3] = { 0.0, 0.0, 0.0 };
                if( c < w )
                {
                    const double * col = Ls + (int64_t)c * r;
                    #pragma unroll 8
                    for( int a = c + gl; a < w; a += G2 )
                    {
                        const double l = col[a];
                        for( int k = 0; k < 3; ++k ) { acc[k] += l * t[3 * a + k]; }
                    }
                }
                for( int k = 0; k < 3; ++k ) { acc[k] = group_sum( acc[k], G2 ); }
                if( c < w && gl == 0 ) { for( int k = 0; k < 3; ++k ) { F.xp[3 * (int64_t)(f + c) + k] = acc[k]; } }
            }
        }
    }

    // ---------------------------------------------------------------------------------------------
    // Split form for levels with large supernodes: the work of one supernode is spread over many
    // workgroups through task lists (int2 = supernode, first row or column).
    //   forward  1: one workgroup per supernode: init and merge the children; the supernode's own
    //               rows go to the column buffer aux, the update rows stay in z
    //   forward  d: one workgroup per 64 rows of the diagonal block: z[0:w] = inv(L_ss) aux
    //   forward  2: one workgroup per 64 update rows: z[q] -= L(q, :) y
    //               (forward d and 2: lane = row, the 4 waves split the columns, sums meet in LDS)
    //   backward 1: one workgroup per group of columns: t[c] = y[c] - L(:, c)^T x(R[w:]) (in place)
    //   backward 2: one workgroup per group of columns: x[f + c] = inv(L_ss)(:, c)^T t
    constexpr int CHOL_SPLIT_BS = 256;

    template<int BS>
    __global__ __launch_bounds__(BS) void k_chol_fwd1( CholDev F, const int32_t * nodes, double alpha, const double * B, int64_t ldB )
    {
        extern __shared__ double sh[];
        const int s = nodes[blockIdx.x];
        const int f = F.sn_begin[s];
        const int w = F.sn_begin[s + 1] - f;
        const int64_t rb = F.row_ptr[s];
        const int r = (int)( F.row_ptr[s + 1] - rb );
        double * z = F.z + 3 * rb;
        const double * Ls = F.vals + F.val_ptr[s];
        double * zs = sh;            // 3 w
        const int tid = threadIdx.x;
        for( int a = tid; a < w; a += BS )
        {
            const int64_t o = F.perm[f + a];
            for( int k = 0; k < 3; ++k ) { zs[3 * a + k] = alpha * B[ldB * o + k]; }
        }
        for( int q = w + tid; q < r; q += BS ) { for( int k = 0; k < 3; ++k ) { z[3 * q + k] = 0.0; } }
        __syncthreads();
        for( int kk = F.kid_ptr[s]; kk < F.kid_ptr[s + 1]; ++kk )
        {
            const int c = F.kids[kk];
            const int wc = F.sn_begin[c + 1] - F.sn_begin[c];
            const int64_t rbc = F.row_ptr[c];
            const int rc = (int)( F.row_ptr[c + 1] - rbc );
            const double * zc = F.z + 3 * rbc;
            for( int q = wc + tid; q < rc; q += BS )
            {
                const int pos = F.cmap[rbc + q];
                if( pos < w ) { for( int k = 0; k < 3; ++k ) { zs[3 * pos + k] += zc[3 * q + k]; } }
                else          { for( int k = 0; k < 3; ++k ) { z[3 * pos + k]  += zc[3 * q + k]; } }
            }
            __syncthreads();
        }
        (void)Ls;
        double * aux = F.aux + 3 * (int64_t)f;
        for( int a = tid; a < 3 * w; a += BS ) { aux[a] = zs[a]; }
    }

    // Forward, split form, diagonal block: y = Inv * aux (Inv lower triangular, the top w x w of the
    // supernode block) into the front's first w rows. 64 rows per task, one per lane; the 4 waves
    // take a quarter of the columns each and their partial sums meet in LDS.
    __global__ __launch_bounds__(256) void k_chol_fwd_diag( CholDev F, const int2 * tasks )
    {
        extern __shared__ double sh[];
        double * xs = sh;                       // 3 w
        double * part = sh + 3 * CHOL_MAX_WIDTH; // 4 x 64 x 3
        const int2 tk = tasks[blockIdx.x];
        const int s = tk.x;
        const int f = F.sn_begin[s];
        const int w = F.sn_begin[s + 1] - f;
        const int64_t rb = F.row_ptr[s];
        const int r = (int)( F.row_ptr[s + 1] - rb );
        const double * Ls = F.vals + F.val_ptr[s];
        const double * aux = F.aux + 3 * (int64_t)f;
        for( int a = threadIdx.x; a < 3 * w; a += 256 ) { xs[a] = aux[a]; }
        __syncthreads();
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        const int a = tk.y + lane;
        const int cw = ( w + 3 ) / 4;
        double acc[3] = { 0.0, 0.0, 0.0 };
        if( a < w )
        {
            const int cb = wv * cw, ce = min( a + 1, min( w, cb + cw ) );
            #pragma unroll 8
            for( int c = cb; c < ce; ++c )
            {
                const double l = Ls[(int64_t)c * r + a];
                for( int k = 0; k < 3; ++k ) { acc[k] += l * xs[3 * c + k]; }
            }
        }
        for( int k = 0; k < 3; ++k ) { part[(wv * 64 + lane) * 3 + k] = acc[k]; }
        __syncthreads();
        if( wv == 0 && a < w )
        {
            double * z = F.z + 3 * rb;
            for( int k = 0; k < 3; ++k )
            {
                z[3 * a + k] = ( part[lane * 3 + k] + part[(64 + lane) * 3 + k] ) + ( part[(128 + lane) * 3 + k] + part[(192 + lane) * 3 + k] );
            }
        }
    }

    // Forward, split form, update rows: z[q] -= L(q, 0:w) y for 64 rows q per task (one per lane),
    // the 4 waves taking a quarter of the columns each, partial sums added in LDS.
    __global__ __launch_bounds__(256) void k_chol_fwd2( CholDev F, const int2 * tasks )
    {
        extern __shared__ double sh[];
        double * ys = sh;                        // 3 w
        double * part = sh + 3 * CHOL_MAX_WIDTH; // 4 x 64 x 3
        const int2 tk = tasks[blockIdx.x];
        const int s = tk.x;
        const int w = F.sn_begin[s + 1] - F.sn_begin[s];
        const int64_t rb = F.row_ptr[s];
        const int r = (int)( F.row_ptr[s + 1] - rb );
        double * z = F.z + 3 * rb;
        const double * Ls = F.vals + F.val_ptr[s];
        for( int a = threadIdx.x; a < 3 * w; a += 256 ) { ys[a] = z[a]; }
        __syncthreads();
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        const int q = tk.y + lane;
        const int cw = ( w + 3 ) / 4;
        double acc[3] = { 0.0, 0.0, 0.0 };
        if( q < r )
        {
            const int cb = wv * cw, ce = min( w, cb + cw );
            #pragma unroll 8
            for( int c = cb; c < ce; ++c )
            {
                const double l = Ls[(int64_t)c * r + q];
                for( int k = 0; k < 3; ++k ) { acc[k] += l * ys[3 * c + k]; }
            }
        }
        for( int k = 0; k < 3; ++k ) { part[(wv * 64 + lane) * 3 + k] = acc[k]; }
        __syncthreads();
        if( wv == 0 && q < r )
        {
            for( int k = 0; k < 3; ++k )
            {
                z[3 * q + k] -= ( part[lane * 3 + k] + part[(64 + lane) * 3 + k] ) + ( part[(128 + lane) * 3 + k] + part[(192 + lane) * 3 + k] );
            }
        }
    }

    // group size for reductions of length n: a power of two, at most 64
    __device__ __forceinline__ int chol_group( int n ) { return pow2_at_least( n < 64 ? n : 64 ); }

    __global__ __launch_bounds__(CHOL_SPLIT_BS) void k_chol_bwd1( CholDev F, const int2 * tasks )
    {
        const int2 tk = tasks[blockIdx.x];
        const int s = tk.x;
        const int w = F.sn_begin[s + 1] - F.sn_begin[s];
        const int64_t rb = F.row_ptr[s];
        const int r = (int)( F.row_ptr[s + 1] - rb );
        const int u = r - w;
        double * z = F.z + 3 * rb;
        const int32_t * R = F.rows + rb;
        const double * Ls = F.vals + F.val_ptr[s];
        const int G = chol_group( u );
        const int grp = threadIdx.x / G, gl = threadIdx.x % G;
        const int c = tk.y + grp;
        double acc[3] = { 0.0, 0.0, 0.0 };
        if( c < w )
        {
            const double * col = Ls + (int64_t)c * r;
            #pragma unroll 8
            for( int q = w + gl; q < r; q += G )
            {
                const double l = col[q];
                const int64_t j = R[q];
                for( int k = 0; k < 3; ++k ) { acc[k] += l * F.xp[3 * j + k]; }
            }
        }
        for( int k = 0; k < 3; ++k ) { acc[k] = group_sum( acc[k], G ); }
        if( c < w && gl == 0 ) { for( int k = 0; k < 3; ++k ) { z[3 * c + k] -= acc[k]; } }
    }

    __global__ __launch_bounds__(CHOL_SPLIT_BS) void k_chol_bwd2( CholDev F, const int2 * tasks )
    {
        const int2 tk = tasks[blockIdx.x];
        const int s = tk.x;
        const int f = F.sn_begin[s];
        const int w = F.sn_begin[s + 1] - f;
        const int64_t rb = F.row_ptr[s];
        const int r = (int)( F.row_ptr[s + 1] - rb );
        const double * z = F.z + 3 * rb;
        const double * Ls = F.vals + F.val_ptr[s];
        const int G = chol_group( w );
        const int grp = threadIdx.x / G, gl = threadIdx.x % G;
        const int c = tk.y + grp;
        double acc[3] = { 0.0, 0.0, 0.0 };
        if( c < w )
        {
            const double * col = Ls + (int64_t)c * r;
            #pragma unroll 8
            for( int a = c + gl; a < w; a += G )
            {
                const double l = col[a];
                for( int k = 0; k < 3; ++k ) { acc[k] += l * z[3 * a + k]; }
            }
        }
        for( int k = 0; k < 3; ++k ) { acc[k] = group_sum( acc[k], G ); }
        if( c < w && gl == 0 ) { for( int k = 0; k < 3; ++k ) { F.xp[3 * (int64_t)(f + c) + k] = acc[k]; } }
    }

    __global__ void k_chol_scatter( int64_t n, const int32_t * perm, const double * xp, double beta, double * Y, int64_t ldY )
    {
        const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if( j >= n ) { return; }
        double * y = Y + ldY * (int64_t)perm[j];
        for( int k = 0; k < 3; ++k )
        {
            const double v = xp[3 * j + k];
            y[k] = (beta == 0.0) ? v : v + beta * y[k];
        }
    }

    void launch_chol_fwd_level( const CholDev & F, const int32_t * nodes, int64_t count, int wmax, bool big,
                                double alpha, const double * B, int64_t ldB, hipStream_t s )
    {
        if( count <= 0 ) { return; }
        const size_t shm = sizeof(double) * 6 * (size_t)wmax;
        if( big ) { hipLaunchKernelGGL( (k_chol_fwd<256>), dim3((unsigned)count), dim3(256), shm, s, F, nodes, alpha, B, ldB ); }
        else      { hipLaunchKernelGGL( (k_chol_fwd<64>),  dim3((unsigned)count), dim3(64),  shm, s, F, nodes, alpha, B, ldB ); }
    }

    void launch_chol_bwd_level( const CholDev & F, const int32_t * nodes, int64_t count, int wmax, bool big, hipStream_t s )
    {
        if( count <= 0 ) { return; }
        const size_t shm = sizeof(double) * ( 3 * (size_t)wmax + 3 * (size_t)CHOL_QC );
        if( big ) { hipLaunchKernelGGL( (k_chol_bwd<256>), dim3((unsigned)count), dim3(256), shm, s, F, nodes ); }
        else      { hipLaunchKernelGGL( (k_chol_bwd<64>),  dim3((unsigned)count), dim3(64),  shm, s, F, nodes ); }
    }

    void launch_chol_fwd1_level( const CholDev & F, const int32_t * nodes, int64_t count, int wmax,
                                 double alpha, const double * B, int64_t ldB, hipStream_t s )
    {
        if( count <= 0 ) { return; }
        const size_t shm = sizeof(double) * 3 * (size_t)wmax;
        hipLaunchKernelGGL( (k_chol_fwd1<CHOL_SPLIT_BS>), dim3((unsigned)count), dim3(CHOL_SPLIT_BS), shm, s, F, nodes, alpha, B, ldB );
    }

    void launch_chol_tasks( int kind, const CholDev & F, const int2 * tasks, int64_t count, int wmax, hipStream_t s )
    {
        if( count <= 0 ) { return; }
        const dim3 grid( (unsigned)count ), block( CHOL_SPLIT_BS );
        const size_t fwd_shm = sizeof(double) * ( 3 * (size_t)CHOL_MAX_WIDTH + 4 * 64 * 3 );
        (void)wmax;
        if( kind == 0 )      { hipLaunchKernelGGL( k_chol_fwd2, grid, block, fwd_shm, s, F, tasks ); }
        else if( kind == 1 ) { hipLaunchKernelGGL( k_chol_bwd1, grid, block, 0, s, F, tasks ); }
        else if( kind == 2 ) { hipLaunchKernelGGL( k_chol_bwd2, grid, block, 0, s, F, tasks ); }
        else                 { hipLaunchKernelGGL( k_chol_fwd_diag, grid, block, fwd_shm, s, F, tasks ); }
    }

    void launch_chol_scatter( int64_t n, const int32_t * perm, const double * xp, double beta, double * Y, int64_t ldY, hipStream_t s )
    {
        if( n <= 0 ) { return; }
        hipLaunchKernelGGL( k_chol_scatter, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, perm, xp, beta, Y, ldY );
    }

} // namespace rp
