// Pseudo-Laplacian L_{s'} of the reference preconditioner (PseudoLaplacian<Mesh,false>, s' = 2 - s):
// scalar metric blocks val_ij = -(r_ij^2)^{s_exp}, s_exp = -n/2 - s', on the near field (primitive
// barycenters) and the far field (cluster centers), diagonal diag_i = -sum_j (a_j/a_i) val_ij, and
// the block product on 3 right-hand sides with the low-order pre/post operator
// (P_in_i = a_i mean(X over simplex i)).
//
// Reference: src/Metrics/PseudoLaplacian/PseudoLaplacian_Kernel_NF.hpp:124-146,
// PseudoLaplacian_Kernel_FF.hpp:101-146, PseudoLaplacian_Traversor.hpp:143-300,
// src/Metrics/PseudoLaplacian.hpp:16-107 (LowOrder pre/post, ClusterTree.hpp:770-943).
// Same leaf-tile traversal and chunk staging as the tangent-point kernels (k_nf.hip).

#include "k_pair.hpp"
#include "rp_launch.hpp"

#include <hip/hip_bf16.h>
#include <type_traits>

namespace rp
{
    constexpr int PL_CH = 64;

    // near-field value storage of the pseudo-Laplacian: FP64, FP32, or BF16 (the value computed in FP32
    // and rounded to nearest even; the diagonal is summed from the rounded values, so the stored operator
    // is the pseudo-Laplacian of the rounded weights and annihilates constants as the exact one does)
    constexpr int PL_F64 = 0, PL_F32 = 1, PL_BF16 = 2;
    __device__ __forceinline__ double pl_val( double x ) { return x; }
    __device__ __forceinline__ double pl_val( float x ) { return (double)x; }
    __device__ __forceinline__ double pl_val( __hip_bfloat16 x ) { return (double)__bfloat162float( x ); }

    // ------------------------------------------------------------------ near-field assembly
    template<int VM>
    __global__ __launch_bounds__(64) void k_nf_pl_assemble( NFArgs A )
    {
        using R = typename std::conditional<VM == PL_F64, double, float>::type;
        __shared__ double sy[2][4][PL_CH];   // b, y0, y1, y2
        __shared__ uint32_t sm[2][PL_CH];
        __shared__ double red[64];
        const int lane = threadIdx.x;
        const int64_t I = xcd_leaf( A.L );
        if( I < 0 ) { return; }
        const int ib = A.leaf_begin[I];
        const int nI = A.leaf_end[I] - ib;
        const int nS = 64 / nI;
        const int ii = lane % nI;
        const int c  = lane / nI;
        const bool active = lane < nS * nI;
        const int64_t i = ib + ii;
        const int64_t T = A.T;
        const double a = A.a[i];
        const double x0 = A.x[i], x1 = A.x[T + i], x2 = A.x[2 * T + i];
        const double inv_a = 1.0 / a;
        const int64_t cb = A.col_ptr[I];
        const int64_t M_ = A.col_ptr[I + 1] - cb;
        const int64_t vb = A.val_ptr[I];
        double dacc = 0.0;
        double rd[4];
        uint32_t rm = 0;
        auto gload = [&]( int64_t base )
        {
            const int64_t m = base + lane;
            if( m < M_ )
            {
                const int32_t j = A.col_prim[cb + m];
                rm = A.col_mask[cb + m];
                const double2 * q = reinterpret_cast<const double2 *>( A.rec + 10 * (int64_t)j );
                const double2 q0 = q[0], q1 = q[1];
                rd[0] = q0.x; rd[1] = q0.y; rd[2] = q1.x; rd[3] = q1.y;
            }
            else { rm = 0u; }
        };
        auto lstore = [&]( int buf ) { sm[buf][lane] = rm; for( int k = 0; k < 4; ++k ) { sy[buf][k][lane] = rd[k]; } };
        gload( 0 );
        lstore( 0 );
        __syncthreads();
        int buf = 0;
        for( int64_t base = 0; base < M_; base += PL_CH )
        {
            const bool more = base + PL_CH < M_;
            if( more ) { gload( base + PL_CH ); }
            const int cnt_c = (int)( (M_ - base < PL_CH) ? (M_ - base) : PL_CH );
            if( active )
            {
                for( int mm = c; mm < cnt_c; mm += nS )
                {
                    const int64_t slot = vb + (base + mm) * nI + ii;
                    if( (sm[buf][mm] >> ii) & 1u )
                    {
                        const double b = sy[buf][0][mm];
                        const R v0 = (R)(sy[buf][1][mm] - x0), v1 = (R)(sy[buf][2][mm] - x1), v2 = (R)(sy[buf][3][mm] - x2);
                        const R r2 = v0 * v0 + v1 * v1 + v2 * v2;
                        const R val = -metric_power<R>( A.ex, r2 );
                        double stored;
                        if constexpr ( VM == PL_BF16 )
                        {
                            const __hip_bfloat16 h = __float2bfloat16( val );
                            reinterpret_cast<__hip_bfloat16 *>(A.vals)[slot] = h;
                            stored = pl_val( h );
                        }
                        else if constexpr ( VM == PL_F32 ) { reinterpret_cast<float *>(A.vals)[slot] = val; stored = (double)val; }
                        else                               { reinterpret_cast<double *>(A.vals)[slot] = val; stored = val; }
                        dacc -= ( b * inv_a ) * stored;
                    }
                    else
                    {
                        if constexpr ( VM == PL_BF16 )     { reinterpret_cast<__hip_bfloat16 *>(A.vals)[slot] = __float2bfloat16( 0.0f ); }
                        else if constexpr ( VM == PL_F32 ) { reinterpret_cast<float *>(A.vals)[slot] = 0.0f; }
                        else                               { reinterpret_cast<double *>(A.vals)[slot] = 0.0; }
                    }
                }
            }
            if( more ) { lstore( buf ^ 1 ); }
            __syncthreads();
            buf ^= 1;
        }
        red[lane] = dacc;
        __syncthreads();
        if( lane < nI )
        {
            double s = 0.0;
            for( int cc = 0; cc < nS; ++cc ) { s += red[lane + cc * nI]; }
            A.diag[i] = s;
        }
    }

    // ------------------------------------------------------------------ far-field assembly
    template<bool MIXED>
    __global__ __launch_bounds__(256) void k_ff_pl_assemble( FFArgs A )
    {
        using R = typename std::conditional<MIXED, float, double>::type;
        const int w = threadIdx.x >> 6;
        const int lane = threadIdx.x & 63;
        const int64_t i = (int64_t)blockIdx.x * 4 + w;
        if( i >= A.C ) { return; }
        const double inv_a = 1.0 / A.a[i];
        const double x0 = A.x[i], x1 = A.x[A.C + i], x2 = A.x[2 * A.C + i];
        double dacc = 0.0;
        for( int64_t k = A.rp[i] + lane; k < A.rp[i + 1]; k += 64 )
        {
            const int32_t j = A.ci[k];
            const R v0 = (R)(A.x[j] - x0), v1 = (R)(A.x[A.C + j] - x1), v2 = (R)(A.x[2 * A.C + j] - x2);
            const R r2 = v0 * v0 + v1 * v1 + v2 * v2;
            const R val = -metric_power<R>( A.ex, r2 );
            if constexpr ( MIXED ) { reinterpret_cast<float *>(A.vals)[k] = val; }
            else                   { reinterpret_cast<double *>(A.vals)[k] = val; }
            dacc -= ( A.a[j] * inv_a ) * (double)val;
        }
        #pragma unroll
        for( int off = 32; off > 0; off >>= 1 ) { dacc += __shfl_xor( dacc, off, 64 ); }
        if( lane == 0 ) { A.diag[i] = dacc; }
    }

    // ------------------------------------------------------------------ low-order pre-operator
    template<int SIZE>
    __global__ void k_pl_pre( int64_t T, const int32_t * cols, const double * pre_vals, const double * X, int64_t ldX, double * P_in )
    {
        const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if( i >= T ) { return; }
        double s[3] = { 0, 0, 0 };
        for( int l = 0; l < SIZE; ++l )
        {
            const double wgt = pre_vals[4 * SIZE * i + l];
            const double * x = X + ldX * (int64_t)cols[SIZE * i + l];
            for( int k = 0; k < 3; ++k ) { s[k] += wgt * x[k]; }
        }
        for( int k = 0; k < 3; ++k ) { P_in[3 * i + k] = s[k]; }
    }

    // ------------------------------------------------------------------ near-field product + post epilogue
    // (the occupancy hint of 8 wavefronts per SIMD measured 7 us faster per preconditioner application
    // than none, in repeated runs, the same sums; 50-51 registers, no spills)
    template<typename VT, int SIZE>
    __global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8))) void k_nf_pl_spmm( int64_t L, const int32_t * leaf_begin, const int32_t * leaf_end,
        const int64_t * col_ptr, const int32_t * col_prim, const uint32_t * col_mask, const int64_t * val_ptr,
        const VT * vals, const double * diag, const double * xin, const double * leaf_out, int use_far,
        const double * pre_vals, double * S_out )
    {
        __shared__ double sx[2][3][PL_CH];
        __shared__ double red[64][3];
        const int lane = threadIdx.x;
        const int64_t I = xcd_leaf( L );
        if( I < 0 ) { return; }
        const int ib = leaf_begin[I];
        const int nI = leaf_end[I] - ib;
        const int nS = 64 / nI;
        const int ii = lane % nI;
        const int c  = lane / nI;
        const bool active = lane < nS * nI;
        const int64_t cb = col_ptr[I];
        const int64_t M_ = col_ptr[I + 1] - cb;
        const int64_t vb = val_ptr[I];
        double acc[3] = { 0, 0, 0 };
        double rx[3] = { 0, 0, 0 };
        // column indices are fetched a chunk before the inputs they address
        auto iload = [&]( int64_t base ) -> int32_t
        {
            const int64_t m = base + lane;
            return ( m < M_ ) ? col_prim[cb + m] : -1;
        };
        auto gload = [&]( int32_t j )
        {
            if( j >= 0 ) { for( int k = 0; k < 3; ++k ) { rx[k] = xin[3 * (int64_t)j + k]; } }
        };
        auto lstore = [&]( int buf ) { for( int k = 0; k < 3; ++k ) { sx[buf][k][lane] = rx[k]; } };
        int32_t j_next = iload( PL_CH );
        gload( iload( 0 ) );
        lstore( 0 );
        __syncthreads();
        int buf = 0;
        for( int64_t base = 0; base < M_; base += PL_CH )
        {
            const bool more = base + PL_CH < M_;
            if( more )
            {
                gload( j_next );
                j_next = iload( base + 2 * PL_CH );
            }
            const int cnt_c = (int)( (M_ - base < PL_CH) ? (M_ - base) : PL_CH );
            if( active )
            {
                // all of the lane's values of the chunk (UV = 16 slots for 2- and 4-byte values: one batch
                // when nS >= 4, i.e. leaves of at most 16 primitives) are loaded before the products, so
                // the chunk waits for one round trip; every tile slot exists and the assembly stores 0 in
                // the slots of masked pairs, so the products need no mask; the products are summed in
                // column order, one at a time. Value offsets are 32-bit from the chunk's wave-uniform base.
                constexpr int UV = sizeof(VT) == 8 ? 8 : 16;
                const char * vchunk = reinterpret_cast<const char *>( vals + vb + base * nI );
                const uint32_t ostep = (uint32_t)( nS * nI ) * (uint32_t)sizeof(VT);
                for( int mm = c; mm < cnt_c; mm += UV * nS )
                {
                    VT v[UV];
                    uint32_t ob = (uint32_t)( mm * nI + ii ) * (uint32_t)sizeof(VT);
                    #pragma unroll
                    for( int u = 0; u < UV; ++u )
                    {
                        if( mm + u * nS < cnt_c ) { v[u] = *reinterpret_cast<const VT *>( vchunk + ob ); }
                        ob += ostep;
                    }
                    #pragma unroll
                    for( int u = 0; u < UV; ++u )
                    {
                        const int m2 = mm + u * nS;
                        if( m2 >= cnt_c ) { break; }
                        const double vv = pl_val( v[u] );
                        for( int k = 0; k < 3; ++k ) { acc[k] += vv * sx[buf][k][m2]; }
                    }
                }
            }
            if( more ) { lstore( buf ^ 1 ); }
            __syncthreads();
            buf ^= 1;
        }
        for( int k = 0; k < 3; ++k ) { red[lane][k] = acc[k]; }
        __syncthreads();
        if( lane < nI )
        {
            const int64_t i = ib + ii;
            double s[3] = { 0, 0, 0 };
            for( int cc = 0; cc < nS; ++cc ) { for( int k = 0; k < 3; ++k ) { s[k] += red[lane + cc * nI][k]; } }
            const double d = diag[i];
            for( int k = 0; k < 3; ++k ) { s[k] += d * xin[3 * i + k]; }
            if( use_far ) { for( int k = 0; k < 3; ++k ) { s[k] = s[k] + leaf_out[3 * I + k]; } }
            for( int sl = 0; sl < SIZE; ++sl )
            {
                const double wgt = pre_vals[4 * SIZE * i + sl];
                for( int k = 0; k < 3; ++k ) { S_out[(SIZE * i + sl) * 3 + k] = wgt * s[k]; }
            }
        }
    }

    template<int SIZE>
    __global__ void k_pl_post_prim( int64_t T, const int32_t * leaf_idx_of_prim, const double * pre_vals,
                                    const double * leaf_out, int use_far, double * S_out )
    {
        const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if( i >= T ) { return; }
        double s[3];
        for( int k = 0; k < 3; ++k ) { s[k] = use_far ? leaf_out[3 * (int64_t)leaf_idx_of_prim[i] + k] : 0.0; }
        for( int sl = 0; sl < SIZE; ++sl )
        {
            const double wgt = pre_vals[4 * SIZE * i + sl];
            for( int k = 0; k < 3; ++k ) { S_out[(SIZE * i + sl) * 3 + k] = wgt * s[k]; }
        }
    }

    template<typename VT>
    __global__ __launch_bounds__(256) void k_ff_pl_spmm( int64_t C, const int64_t * rp, const int32_t * ci,
        const VT * vals, const double * diag, const double * xin, double * yout )
    {
        constexpr int G = 16;   // lanes per cluster row (rows hold about two dozen entries)
        const int lane = threadIdx.x % G;
        const int64_t i = ( (int64_t)blockIdx.x * 256 + threadIdx.x ) / G;
        if( i >= C ) { return; }
        double acc[3] = { 0, 0, 0 };
        // entries k, k + G, ... in that order, two per step with both loaded before either product
        const int64_t e = rp[i + 1];
        for( int64_t k = rp[i] + lane; k < e; k += 2 * G )
        {
            const bool two = k + G < e;
            const int32_t j0 = ci[k];
            const double v0 = (double)vals[k];
            int32_t j1 = j0;
            double v1 = 0.0;
            if( two ) { j1 = ci[k + G]; v1 = (double)vals[k + G]; }
            const double * x0 = xin + 3 * (int64_t)j0;
            const double * x1 = xin + 3 * (int64_t)j1;
            double y0[3], y1[3];
            for( int h = 0; h < 3; ++h ) { y0[h] = x0[h]; y1[h] = x1[h]; }
            for( int h = 0; h < 3; ++h ) { acc[h] += v0 * y0[h]; }
            if( two ) { for( int h = 0; h < 3; ++h ) { acc[h] += v1 * y1[h]; } }
        }
        for( int h = 0; h < 3; ++h )
        {
            #pragma unroll
            for( int off = G / 2; off > 0; off >>= 1 ) { acc[h] += __shfl_xor( acc[h], off, 64 ); }
        }
        if( lane == 0 )
        {
            const double d = diag[i];
            for( int h = 0; h < 3; ++h ) { yout[3 * i + h] = acc[h] + d * xin[3 * i + h]; }
        }
    }

    // ------------------------------------------------------------------ launchers
    void launch_pl_assemble( int precision, const NFArgs & A, const FFArgs & B, bool do_nf, bool do_ff, hipStream_t s )
    {
        const bool mixed = precision != RP_FP64;   // the far field: FP32 unless FP64
        if( do_nf && A.L )
        {
            if( precision == RP_PL_BF16 ) { hipLaunchKernelGGL( (k_nf_pl_assemble<PL_BF16>), leaf_grid( A.L ), dim3(64), 0, s, A ); }
            else if( mixed )              { hipLaunchKernelGGL( (k_nf_pl_assemble<PL_F32>),  leaf_grid( A.L ), dim3(64), 0, s, A ); }
            else                          { hipLaunchKernelGGL( (k_nf_pl_assemble<PL_F64>),  leaf_grid( A.L ), dim3(64), 0, s, A ); }
        }
        if( do_ff && B.C )
        {
            const dim3 grid( (unsigned)((B.C + 3) / 4) );
            if( mixed ) { hipLaunchKernelGGL( (k_ff_pl_assemble<true>), grid, dim3(256), 0, s, B ); }
            else        { hipLaunchKernelGGL( (k_ff_pl_assemble<false>), grid, dim3(256), 0, s, B ); }
        }
    }

    void launch_pl_pre( int size, int64_t T, const int32_t * cols, const double * pre_vals, const double * X, int64_t ldX, double * P_in, hipStream_t s )
    {
        if( !T ) { return; }
        const dim3 grid( (unsigned)((T + 255) / 256) );
        if( size == 3 ) { hipLaunchKernelGGL( (k_pl_pre<3>), grid, dim3(256), 0, s, T, cols, pre_vals, X, ldX, P_in ); }
        else            { hipLaunchKernelGGL( (k_pl_pre<2>), grid, dim3(256), 0, s, T, cols, pre_vals, X, ldX, P_in ); }
    }

    void launch_pl_nf_spmm( int precision, int size, int64_t L, const int32_t * lb, const int32_t * le, const int64_t * cp,
                            const int32_t * cpr, const uint32_t * cm, const int64_t * vp, const void * vals, const double * diag,
                            const double * xin, const double * leaf_out, int use_far, const double * pre_vals, double * S_out, hipStream_t s )
    {
        if( !L ) { return; }
        const dim3 grid = leaf_grid( L ), block( 64 );
        if( size == 3 )
        {
            if( precision == RP_PL_BF16 ) { hipLaunchKernelGGL( (k_nf_pl_spmm<__hip_bfloat16,3>), grid, block, 0, s, L, lb, le, cp, cpr, cm, vp, (const __hip_bfloat16 *)vals, diag, xin, leaf_out, use_far, pre_vals, S_out ); }
            else if( precision == RP_MIXED ) { hipLaunchKernelGGL( (k_nf_pl_spmm<float,3>), grid, block, 0, s, L, lb, le, cp, cpr, cm, vp, (const float *)vals, diag, xin, leaf_out, use_far, pre_vals, S_out ); }
            else        { hipLaunchKernelGGL( (k_nf_pl_spmm<double,3>), grid, block, 0, s, L, lb, le, cp, cpr, cm, vp, (const double *)vals, diag, xin, leaf_out, use_far, pre_vals, S_out ); }
        }
        else
        {
            if( precision == RP_PL_BF16 ) { hipLaunchKernelGGL( (k_nf_pl_spmm<__hip_bfloat16,2>), grid, block, 0, s, L, lb, le, cp, cpr, cm, vp, (const __hip_bfloat16 *)vals, diag, xin, leaf_out, use_far, pre_vals, S_out ); }
            else if( precision == RP_MIXED ) { hipLaunchKernelGGL( (k_nf_pl_spmm<float,2>), grid, block, 0, s, L, lb, le, cp, cpr, cm, vp, (const float *)vals, diag, xin, leaf_out, use_far, pre_vals, S_out ); }
            else        { hipLaunchKernelGGL( (k_nf_pl_spmm<double,2>), grid, block, 0, s, L, lb, le, cp, cpr, cm, vp, (const double *)vals, diag, xin, leaf_out, use_far, pre_vals, S_out ); }
        }
    }

    void launch_pl_post_prim( int size, int64_t T, const int32_t * leaf_idx_of_prim, const double * pre_vals, const double * leaf_out, int use_far, double * S_out, hipStream_t s )
    {
        if( !T ) { return; }
        const dim3 grid( (unsigned)((T + 255) / 256) );
        if( size == 3 ) { hipLaunchKernelGGL( (k_pl_post_prim<3>), grid, dim3(256), 0, s, T, leaf_idx_of_prim, pre_vals, leaf_out, use_far, S_out ); }
        else            { hipLaunchKernelGGL( (k_pl_post_prim<2>), grid, dim3(256), 0, s, T, leaf_idx_of_prim, pre_vals, leaf_out, use_far, S_out ); }
    }

    void launch_pl_ff_spmm( bool mixed, int64_t C, const int64_t * rp, const int32_t * ci, const void * vals, const double * diag, const double * xin, double * yout, hipStream_t s )
    {
        if( !C ) { return; }
        const dim3 grid( (unsigned)((C + 15) / 16) );
        if( mixed ) { hipLaunchKernelGGL( (k_ff_pl_spmm<float>), grid, dim3(256), 0, s, C, rp, ci, (const float *)vals, diag, xin, yout ); }
        else        { hipLaunchKernelGGL( (k_ff_pl_spmm<double>), grid, dim3(256), 0, s, C, rp, ci, (const double *)vals, diag, xin, yout ); }
    }

} // namespace rp
