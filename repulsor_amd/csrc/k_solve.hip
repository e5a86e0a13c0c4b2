// Vector kernels of the Krylov solver for the metric (three right-hand sides, row-major n x 3).
// Inner products are two-level reductions with a fixed block count and a fixed summation order, so
// a solve is bitwise reproducible. The per-column scalars (rho, p.Ap, |r|^2, active flags) stay in
// device memory; the host reads only |r|^2 once per iteration for the stopping test.

#include "rp_launch.hpp"

namespace rp
{
    constexpr int DOT_BLOCKS = 1024;

    __device__ __forceinline__ void block_sum3( double * acc, double (*red)[3] )
    {
        const int tid = threadIdx.x;
        for( int k = 0; k < 3; ++k ) { red[tid][k] = acc[k]; }
        __syncthreads();
        for( int h = blockDim.x / 2; h > 0; h >>= 1 )
        {
            if( tid < h ) { for( int k = 0; k < 3; ++k ) { red[tid][k] += red[tid + h][k]; } }
            __syncthreads();
        }
    }

    __global__ __launch_bounds__(256) void k_dot3_partial( int64_t n, const double * u, const double * v, double * part )
    {
        __shared__ double red[256][3];
        double acc[3] = { 0.0, 0.0, 0.0 };
        for( int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256 )
        {
            for( int k = 0; k < 3; ++k ) { acc[k] += u[3 * i + k] * v[3 * i + k]; }
        }
        block_sum3( acc, red );
        if( threadIdx.x == 0 ) { for( int k = 0; k < 3; ++k ) { part[3 * blockIdx.x + k] = red[0][k]; } }
    }

    __global__ __launch_bounds__(256) void k_dot3_final( const double * part, int nb, double * out )
    {
        __shared__ double red[256][3];
        double acc[3] = { 0.0, 0.0, 0.0 };
        for( int b = threadIdx.x; b < nb; b += 256 ) { for( int k = 0; k < 3; ++k ) { acc[k] += part[3 * b + k]; } }
        block_sum3( acc, red );
        if( threadIdx.x == 0 ) { for( int k = 0; k < 3; ++k ) { out[k] = red[0][k]; } }
    }

    // x += alpha p, r -= alpha Ap with alpha_k = rho_k / pAp_k on active columns; partial sums of |r|^2.
    __global__ __launch_bounds__(256) void k_cg_update( int64_t n, double * x, double * r, const double * p, const double * Ap,
                                                        const double * rho, const double * pAp, const double * active, double * part )
    {
        __shared__ double red[256][3];
        double al[3];
        for( int k = 0; k < 3; ++k ) { al[k] = ( active[k] != 0.0 ) ? rho[k] / pAp[k] : 0.0; }
        double acc[3] = { 0.0, 0.0, 0.0 };
        for( int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256 )
        {
            for( int k = 0; k < 3; ++k )
            {
                x[3 * i + k] += al[k] * p[3 * i + k];
                const double rk = r[3 * i + k] - al[k] * Ap[3 * i + k];
                r[3 * i + k] = rk;
                acc[k] += rk * rk;
            }
        }
        block_sum3( acc, red );
        if( threadIdx.x == 0 ) { for( int k = 0; k < 3; ++k ) { part[3 * blockIdx.x + k] = red[0][k]; } }
    }

    // p = z + beta p with beta_k = rho_new_k / rho_old_k on active columns.
    __global__ __launch_bounds__(256) void k_cg_dir( int64_t n, double * p, const double * z, const double * rho_new, const double * rho_old, const double * active )
    {
        const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
        if( i >= n ) { return; }
        for( int k = 0; k < 3; ++k )
        {
            if( active[k] != 0.0 ) { p[3 * i + k] = z[3 * i + k] + ( rho_new[k] / rho_old[k] ) * p[3 * i + k]; }
        }
    }

    // active_k = |r_k|^2 > tol^2 |b_k|^2
    __global__ void k_cg_check( const double * rr, const double * bb, double tol2, double * active )
    {
        const int k = threadIdx.x;
        if( k < 3 ) { active[k] = ( rr[k] > tol2 * bb[k] ) ? 1.0 : 0.0; }
    }

    __global__ __launch_bounds__(256) void k_copy3( int64_t n, const double * src, int64_t lds, double * dst, int64_t ldd )
    {
        const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
        if( i >= n ) { return; }
        for( int k = 0; k < 3; ++k ) { dst[ldd * i + k] = src[lds * i + k]; }
    }

    // Y = alpha X + beta Y (beta == 0 does not read Y)
    __global__ __launch_bounds__(256) void k_axpby3( int64_t n, double alpha, const double * X, int64_t ldX, double beta, double * Y, int64_t ldY )
    {
        const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
        if( i >= n ) { return; }
        for( int k = 0; k < 3; ++k )
        {
            const double v = alpha * X[ldX * i + k];
            Y[ldY * i + k] = (beta == 0.0) ? v : v + beta * Y[ldY * i + k];
        }
    }

    // ------------------------------------------------------------------ GMRES (Arnoldi with CGS2)
    // Basis vectors are stored back to back (stride 3n). Partial inner products of w with basis
    // vectors 0..nv-1: grid (MDOT_BLOCKS, nv).
    constexpr int MDOT_BLOCKS = 256;

    __global__ __launch_bounds__(256) void k_mdot_partial( int64_t n, const double * Vb, int64_t stride, const double * w, double * part )
    {
        __shared__ double red[256][3];
        const int i = blockIdx.y;
        const double * v = Vb + stride * i;
        double acc[3] = { 0.0, 0.0, 0.0 };
        for( int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < n; r += (int64_t)MDOT_BLOCKS * 256 )
        {
            for( int k = 0; k < 3; ++k ) { acc[k] += v[3 * r + k] * w[3 * r + k]; }
        }
        block_sum3( acc, red );
        if( threadIdx.x == 0 ) { for( int k = 0; k < 3; ++k ) { part[((int64_t)i * MDOT_BLOCKS + blockIdx.x) * 3 + k] = red[0][k]; } }
    }

    // out[i][k] (+)= sum over blocks (fixed order)
    __global__ __launch_bounds__(256) void k_mdot_final( const double * part, double * out, int accumulate )
    {
        __shared__ double red[256][3];
        const int i = blockIdx.x;
        double acc[3] = { 0.0, 0.0, 0.0 };
        for( int b = threadIdx.x; b < MDOT_BLOCKS; b += 256 ) { for( int k = 0; k < 3; ++k ) { acc[k] += part[((int64_t)i * MDOT_BLOCKS + b) * 3 + k]; } }
        block_sum3( acc, red );
        if( threadIdx.x == 0 ) { for( int k = 0; k < 3; ++k ) { out[3 * i + k] = accumulate ? out[3 * i + k] + red[0][k] : red[0][k]; } }
    }

    // w += sign * sum_i coef[i][k] V_i   (per column k)
    __global__ __launch_bounds__(256) void k_maxpy( int64_t n, const double * Vb, int64_t stride, int nv, const double * coef, double sign, double * w )
    {
        const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
        if( r >= n ) { return; }
        double acc[3] = { 0.0, 0.0, 0.0 };
        for( int i = 0; i < nv; ++i )
        {
            const double * v = Vb + stride * i + 3 * r;
            for( int k = 0; k < 3; ++k ) { acc[k] += coef[3 * i + k] * v[k]; }
        }
        for( int k = 0; k < 3; ++k ) { w[3 * r + k] += sign * acc[k]; }
    }

    // out = w / sqrt(nrm2) per column (0 where the norm vanishes)
    __global__ __launch_bounds__(256) void k_normalize3( int64_t n, const double * w, const double * nrm2, double * out )
    {
        const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
        if( r >= n ) { return; }
        for( int k = 0; k < 3; ++k )
        {
            const double s = nrm2[k] > 0.0 ? 1.0 / std::sqrt( nrm2[k] ) : 0.0;
            out[3 * r + k] = w[3 * r + k] * s;
        }
    }

    static dim3 grid256( int64_t n ) { return dim3( (unsigned)std::max<int64_t>( 1, (n + 255) / 256 ) ); }

    // y += alpha x and x *= alpha on plain length-n vectors
    __global__ __launch_bounds__(256) void k_axpy1( int64_t n, double alpha, const double * x, double * y )
    {
        const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
        if( i < n ) { y[i] += alpha * x[i]; }
    }

    __global__ __launch_bounds__(256) void k_scale1( int64_t n, double alpha, double * x )
    {
        const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
        if( i < n ) { x[i] *= alpha; }
    }

    void launch_axpy1( int64_t n, double alpha, const double * x, double * y, hipStream_t s )
    {
        if( n > 0 ) { hipLaunchKernelGGL( k_axpy1, grid256( n ), dim3(256), 0, s, n, alpha, x, y ); }
    }

    void launch_scale1( int64_t n, double alpha, double * x, hipStream_t s )
    {
        if( n > 0 ) { hipLaunchKernelGGL( k_scale1, grid256( n ), dim3(256), 0, s, n, alpha, x ); }
    }

    void launch_mdot( int64_t n, const double * Vb, int64_t stride, int nv, const double * w, double * part, double * out, int accumulate, hipStream_t s )
    {
        if( nv <= 0 ) { return; }
        hipLaunchKernelGGL( k_mdot_partial, dim3(MDOT_BLOCKS, (unsigned)nv), dim3(256), 0, s, n, Vb, stride, w, part );
        hipLaunchKernelGGL( k_mdot_final, dim3((unsigned)nv), dim3(256), 0, s, part, out, accumulate );
    }

    void launch_maxpy( int64_t n, const double * Vb, int64_t stride, int nv, const double * coef, double sign, double * w, hipStream_t s )
    {
        if( n > 0 && nv > 0 ) { hipLaunchKernelGGL( k_maxpy, grid256( n ), dim3(256), 0, s, n, Vb, stride, nv, coef, sign, w ); }
    }

    void launch_normalize3( int64_t n, const double * w, const double * nrm2, double * out, hipStream_t s )
    {
        if( n > 0 ) { hipLaunchKernelGGL( k_normalize3, grid256( n ), dim3(256), 0, s, n, w, nrm2, out ); }
    }

    void launch_dot3( int64_t n, const double * u, const double * v, double * part, double * out, hipStream_t s )
    {
        hipLaunchKernelGGL( k_dot3_partial, dim3(DOT_BLOCKS), dim3(256), 0, s, n, u, v, part );
        hipLaunchKernelGGL( k_dot3_final, dim3(1), dim3(256), 0, s, part, DOT_BLOCKS, out );
    }

    void launch_cg_update( int64_t n, double * x, double * r, const double * p, const double * Ap, const double * rho,
                           const double * pAp, const double * active, double * part, double * rr, hipStream_t s )
    {
        hipLaunchKernelGGL( k_cg_update, dim3(DOT_BLOCKS), dim3(256), 0, s, n, x, r, p, Ap, rho, pAp, active, part );
        hipLaunchKernelGGL( k_dot3_final, dim3(1), dim3(256), 0, s, part, DOT_BLOCKS, rr );
    }

    void launch_cg_dir( int64_t n, double * p, const double * z, const double * rho_new, const double * rho_old, const double * active, hipStream_t s )
    {
        if( n > 0 ) { hipLaunchKernelGGL( k_cg_dir, grid256( n ), dim3(256), 0, s, n, p, z, rho_new, rho_old, active ); }
    }

    void launch_cg_check( const double * rr, const double * bb, double tol2, double * active, hipStream_t s )
    {
        hipLaunchKernelGGL( k_cg_check, dim3(1), dim3(64), 0, s, rr, bb, tol2, active );
    }

    void launch_copy3( int64_t n, const double * src, int64_t lds, double * dst, int64_t ldd, hipStream_t s )
    {
        if( n > 0 ) { hipLaunchKernelGGL( k_copy3, grid256( n ), dim3(256), 0, s, n, src, lds, dst, ldd ); }
    }

    void launch_axpby3( int64_t n, double alpha, const double * X, int64_t ldX, double beta, double * Y, int64_t ldY, hipStream_t s )
    {
        if( n > 0 ) { hipLaunchKernelGGL( k_axpby3, grid256( n ), dim3(256), 0, s, n, alpha, X, ldX, beta, Y, ldY ); }
    }

} // namespace rp
