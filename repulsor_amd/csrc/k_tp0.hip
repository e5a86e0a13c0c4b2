// Near-field (primitive pairs) and far-field (cluster pairs) kernels of TangentPointEnergy0:
// energy, moment derivatives and metric blocks, FP64 and FP32-mixed.
//
// Near field: one wavefront per row leaf I (n_I <= 32 primitives). Lane l handles row
// ii = l mod n_I of the leaf and column slice c = l / n_I of the leaf's concatenated column list
// (all primitives of leaves J in non-admissible blocks with I). Partial derivative sums of the
// slices are combined in LDS in a fixed order, so results are bitwise reproducible. Metric blocks
// are stored per leaf as [column][row] (a0,a1) pairs: the 64 lanes of a wave write consecutive slots.
//
// Far field: one wavefront per cluster row of the full symmetric far pattern; lanes stride the
// row's columns; deterministic shuffle reduction.
//
// Both traverse the full symmetric patterns and accumulate only row-side (S-side) derivatives,
// which equal the reference's DX + DY sums (TP0_Kernel_NF.hpp:245-272) without scatter or atomics.
// The energy counts each unordered pair once (column > row), as the reference's upper-triangle
// traversal does (FMM_Traversor.hpp:100-101).

#include "k_pair.hpp"
#include "rp_launch.hpp"

namespace rp
{
    constexpr int NACC = 13;   // 0: d/da, 1..3: d/d(a x) (scaled), 4..9: d/d(a P), 10,11: diag, 12: energy

    // ------------------------------------------------------------------ near field, FP64
    template<int KIND, bool E, bool D, bool M>
    __global__ __launch_bounds__(256) void k_nf_f64( NFArgs A )
    {
        __shared__ double red[4][64][NACC];
        const int w = threadIdx.x >> 6;
        const int lane = threadIdx.x & 63;
        const int64_t I = (int64_t)blockIdx.x * 4 + w;
        const bool wave_ok = I < A.L;
        double acc[NACC];
        #pragma unroll
        for( int k = 0; k < NACC; ++k ) { acc[k] = 0.0; }
        int nI = 1, nS = 64, ii = 0;
        int64_t i = 0;
        if( wave_ok )
        {
            const int ib = A.leaf_begin[I];
            nI = A.leaf_end[I] - ib;
            nS = 64 / nI;
            ii = lane % nI;
            const int c = lane / nI;
            i = ib + ii;
            if( lane < nS * nI )
            {
                const double a = A.a[i];
                const double x0 = A.x[i], x1 = A.x[A.T + i], x2 = A.x[2 * A.T + i];
                double P[6];
                #pragma unroll
                for( int k = 0; k < 6; ++k ) { P[k] = A.P[k * A.T + i]; }
                const double inv_a = 1.0 / a;
                const int64_t cb = A.col_ptr[I];
                const int64_t M_ = A.col_ptr[I + 1] - cb;
                const int64_t vb = M ? A.val_ptr[I] : 0;
                for( int64_t m = c; m < M_; m += nS )
                {
                    const int32_t j = A.col_prim[cb + m];
                    const uint32_t mk = A.col_mask[cb + m];
                    if( (mk >> ii) & 1u )
                    {
                        const double b = A.a[j];
                        const double v[3] = { A.x[j] - x0, A.x[A.T + j] - x1, A.x[2 * A.T + j] - x2 };
                        double Q[6];
                        #pragma unroll
                        for( int k = 0; k < 6; ++k ) { Q[k] = A.P[k * A.T + j]; }
                        PairTerms<double> t;
                        pair_terms<double,KIND,E,D,M>( A.ex, v, P, Q, t );
                        if constexpr ( E ) { if( j > i ) { acc[12] += a * t.E * b; } }
                        if constexpr ( D )
                        {
                            const double gx = t.g[0] * x0 + t.g[1] * x1 + t.g[2] * x2;
                            acc[0] += b * ( t.E0 + gx );
                            const double sb = A.scale * b;
                            acc[1] -= sb * t.g[0];
                            acc[2] -= sb * t.g[1];
                            acc[3] -= sb * t.g[2];
                            const double bk = b * t.hK;
                            #pragma unroll
                            for( int k = 0; k < 6; ++k ) { acc[4 + k] += bk * t.Vk[k]; }
                        }
                        if constexpr ( M )
                        {
                            reinterpret_cast<double2 *>(A.vals)[vb + m * nI + ii] = make_double2( t.a0, t.a1 );
                            const double boa = b * inv_a;
                            acc[10] -= boa * t.a0;
                            acc[11] -= boa * t.a1;
                        }
                    }
                    else if constexpr ( M )
                    {
                        reinterpret_cast<double2 *>(A.vals)[vb + m * nI + ii] = make_double2( 0.0, 0.0 );
                    }
                }
            }
        }
        #pragma unroll
        for( int k = 0; k < NACC; ++k ) { red[w][lane][k] = acc[k]; }
        __syncthreads();
        if( wave_ok )
        {
            if( lane < nI )
            {
                double s[12];
                #pragma unroll
                for( int k = 0; k < 12; ++k ) { s[k] = 0.0; }
                for( int c = 0; c < nS; ++c )
                {
                    #pragma unroll
                    for( int k = 0; k < 12; ++k ) { s[k] += red[w][lane + c * nI][k]; }
                }
                if constexpr ( D )
                {
                    #pragma unroll
                    for( int k = 0; k < 10; ++k ) { A.dx[k * A.T + i] = s[k]; }
                }
                if constexpr ( M )
                {
                    A.diag[2 * i]     = s[10];
                    A.diag[2 * i + 1] = s[11];
                }
            }
            if constexpr ( E )
            {
                if( lane == 0 )
                {
                    double e = 0.0;
                    for( int l = 0; l < 64; ++l ) { e += red[w][l][12]; }
                    A.epart[I] = e;
                }
            }
        }
    }

    // ------------------------------------------------------------------ near field, FP32 pair arithmetic
    // Coordinates relative to the row leaf's origin o_I (FP64 subtraction, then FP32), pair terms in FP32,
    // FP32 partial sums over chunks of 8 pairs flushed into FP64 accumulators. The area derivative is
    // accumulated in the local frame and moved back with d/da = d/da' - sum_l d/d(a hull_l) . o_I.
    template<int KIND, bool E, bool D, bool M>
    __global__ __launch_bounds__(256) void k_nf_mixed( NFArgs A )
    {
        __shared__ double red[4][64][NACC];
        const int w = threadIdx.x >> 6;
        const int lane = threadIdx.x & 63;
        const int64_t I = (int64_t)blockIdx.x * 4 + w;
        const bool wave_ok = I < A.L;
        double acc[NACC];
        #pragma unroll
        for( int k = 0; k < NACC; ++k ) { acc[k] = 0.0; }
        int nI = 1, nS = 64, ii = 0;
        int64_t i = 0;
        double o0 = 0, o1 = 0, o2 = 0;
        if( wave_ok )
        {
            const int ib = A.leaf_begin[I];
            nI = A.leaf_end[I] - ib;
            nS = 64 / nI;
            ii = lane % nI;
            const int c = lane / nI;
            i = ib + ii;
            o0 = A.origin[I]; o1 = A.origin[A.L + I]; o2 = A.origin[2 * A.L + I];
            if( lane < nS * nI )
            {
                const float a = (float)A.a[i];
                const float x0 = (float)(A.x[i] - o0), x1 = (float)(A.x[A.T + i] - o1), x2 = (float)(A.x[2 * A.T + i] - o2);
                float P[6];
                #pragma unroll
                for( int k = 0; k < 6; ++k ) { P[k] = (float)A.P[k * A.T + i]; }
                const float inv_a = 1.0f / a;
                const float sc = (float)A.scale;
                const int64_t cb = A.col_ptr[I];
                const int64_t M_ = A.col_ptr[I + 1] - cb;
                const int64_t vb = M ? A.val_ptr[I] : 0;
                float part[NACC];
                #pragma unroll
                for( int k = 0; k < NACC; ++k ) { part[k] = 0.0f; }
                int cnt = 0;
                for( int64_t m = c; m < M_; m += nS )
                {
                    const int32_t j = A.col_prim[cb + m];
                    const uint32_t mk = A.col_mask[cb + m];
                    if( (mk >> ii) & 1u )
                    {
                        const float b = (float)A.a[j];
                        const float y0 = (float)(A.x[j] - o0), y1 = (float)(A.x[A.T + j] - o1), y2 = (float)(A.x[2 * A.T + j] - o2);
                        const float v[3] = { y0 - x0, y1 - x1, y2 - x2 };
                        float Q[6];
                        #pragma unroll
                        for( int k = 0; k < 6; ++k ) { Q[k] = (float)A.P[k * A.T + j]; }
                        PairTerms<float> t;
                        pair_terms<float,KIND,E,D,M>( A.ex, v, P, Q, t );
                        if constexpr ( E ) { if( j > i ) { part[12] += a * t.E * b; } }
                        if constexpr ( D )
                        {
                            const float gx = t.g[0] * x0 + t.g[1] * x1 + t.g[2] * x2;
                            part[0] += b * ( t.E0 + gx );
                            const float sb = sc * b;
                            part[1] -= sb * t.g[0];
                            part[2] -= sb * t.g[1];
                            part[3] -= sb * t.g[2];
                            const float bk = b * t.hK;
                            #pragma unroll
                            for( int k = 0; k < 6; ++k ) { part[4 + k] += bk * t.Vk[k]; }
                        }
                        if constexpr ( M )
                        {
                            reinterpret_cast<float2 *>(A.vals)[vb + m * nI + ii] = make_float2( t.a0, t.a1 );
                            const float boa = b * inv_a;
                            part[10] -= boa * t.a0;
                            part[11] -= boa * t.a1;
                        }
                        if( ++cnt == 8 )
                        {
                            #pragma unroll
                            for( int k = 0; k < NACC; ++k ) { acc[k] += (double)part[k]; part[k] = 0.0f; }
                            cnt = 0;
                        }
                    }
                    else if constexpr ( M )
                    {
                        reinterpret_cast<float2 *>(A.vals)[vb + m * nI + ii] = make_float2( 0.0f, 0.0f );
                    }
                }
                #pragma unroll
                for( int k = 0; k < NACC; ++k ) { acc[k] += (double)part[k]; }
            }
        }
        #pragma unroll
        for( int k = 0; k < NACC; ++k ) { red[w][lane][k] = acc[k]; }
        __syncthreads();
        if( wave_ok )
        {
            if( lane < nI )
            {
                double s[12];
                #pragma unroll
                for( int k = 0; k < 12; ++k ) { s[k] = 0.0; }
                for( int c = 0; c < nS; ++c )
                {
                    #pragma unroll
                    for( int k = 0; k < 12; ++k ) { s[k] += red[w][lane + c * nI][k]; }
                }
                if constexpr ( D )
                {
                    // back to the global frame: the hull moment derivative of each of the (n+1) hull
                    // slots equals s[1..3], so sum_l d/d(a hull_l) . o = (n+1) s[1..3] . o
                    s[0] -= (double)A.size * ( s[1] * o0 + s[2] * o1 + s[3] * o2 );
                    #pragma unroll
                    for( int k = 0; k < 10; ++k ) { A.dx[k * A.T + i] = s[k]; }
                }
                if constexpr ( M )
                {
                    A.diag[2 * i]     = s[10];
                    A.diag[2 * i + 1] = s[11];
                }
            }
            if constexpr ( E )
            {
                if( lane == 0 )
                {
                    double e = 0.0;
                    for( int l = 0; l < 64; ++l ) { e += red[w][l][12]; }
                    A.epart[I] = e;
                }
            }
        }
    }

    // ------------------------------------------------------------------ far field
    template<typename T>
    __device__ __forceinline__ T wave_sum( T v )
    {
        #pragma unroll
        for( int off = 32; off > 0; off >>= 1 ) { v += __shfl_xor( v, off, 64 ); }
        return v;
    }

    template<int KIND, bool E, bool D, bool M, bool MIXED>
    __global__ __launch_bounds__(256) void k_ff( FFArgs A )
    {
        const int w = threadIdx.x >> 6;
        const int lane = threadIdx.x & 63;
        const int64_t i = (int64_t)blockIdx.x * 4 + w;
        if( i >= A.C ) { return; }
        double acc[NACC];
        #pragma unroll
        for( int k = 0; k < NACC; ++k ) { acc[k] = 0.0; }
        const double xa = A.a[i];
        const double x0 = A.x[i], x1 = A.x[A.C + i], x2 = A.x[2 * A.C + i];
        const int64_t kb = A.rp[i], ke = A.rp[i + 1];
        if constexpr ( !MIXED )
        {
            double P[6];
            #pragma unroll
            for( int k = 0; k < 6; ++k ) { P[k] = A.P[k * A.C + i]; }
            const double inv_a = 1.0 / xa;
            for( int64_t k = kb + lane; k < ke; k += 64 )
            {
                const int32_t j = A.ci[k];
                const double b = A.a[j];
                const double v[3] = { A.x[j] - x0, A.x[A.C + j] - x1, A.x[2 * A.C + j] - x2 };
                double Q[6];
                #pragma unroll
                for( int h = 0; h < 6; ++h ) { Q[h] = A.P[h * A.C + j]; }
                PairTerms<double> t;
                pair_terms<double,KIND,E,D,M>( A.ex, v, P, Q, t );
                if constexpr ( E ) { if( j > i ) { acc[12] += xa * t.E * b; } }
                if constexpr ( D )
                {
                    const double gx = t.g[0] * x0 + t.g[1] * x1 + t.g[2] * x2;
                    acc[0] += b * ( t.E0 + gx );
                    acc[1] -= b * t.g[0];
                    acc[2] -= b * t.g[1];
                    acc[3] -= b * t.g[2];
                    const double bk = b * t.hK;
                    #pragma unroll
                    for( int h = 0; h < 6; ++h ) { acc[4 + h] += bk * t.Vk[h]; }
                }
                if constexpr ( M )
                {
                    reinterpret_cast<double2 *>(A.vals)[k] = make_double2( t.a0, t.a1 );
                    const double boa = b * inv_a;
                    acc[10] -= boa * t.a0;
                    acc[11] -= boa * t.a1;
                }
            }
        }
        else
        {
            // Local frame centred at the row cluster: x' = 0.
            float P[6];
            #pragma unroll
            for( int k = 0; k < 6; ++k ) { P[k] = (float)A.P[k * A.C + i]; }
            const float af = (float)xa;
            const float inv_a = 1.0f / af;
            for( int64_t k = kb + lane; k < ke; k += 64 )
            {
                const int32_t j = A.ci[k];
                const float b = (float)A.a[j];
                const float v[3] = { (float)(A.x[j] - x0), (float)(A.x[A.C + j] - x1), (float)(A.x[2 * A.C + j] - x2) };
                float Q[6];
                #pragma unroll
                for( int h = 0; h < 6; ++h ) { Q[h] = (float)A.P[h * A.C + j]; }
                PairTerms<float> t;
                pair_terms<float,KIND,E,D,M>( A.ex, v, P, Q, t );
                if constexpr ( E ) { if( j > i ) { acc[12] += (double)( af * t.E * b ); } }
                if constexpr ( D )
                {
                    acc[0] += (double)( b * t.E0 );
                    acc[1] -= (double)( b * t.g[0] );
                    acc[2] -= (double)( b * t.g[1] );
                    acc[3] -= (double)( b * t.g[2] );
                    const float bk = b * t.hK;
                    #pragma unroll
                    for( int h = 0; h < 6; ++h ) { acc[4 + h] += (double)( bk * t.Vk[h] ); }
                }
                if constexpr ( M )
                {
                    reinterpret_cast<float2 *>(A.vals)[k] = make_float2( t.a0, t.a1 );
                    const float boa = b * inv_a;
                    acc[10] -= (double)( boa * t.a0 );
                    acc[11] -= (double)( boa * t.a1 );
                }
            }
        }
        #pragma unroll
        for( int k = 0; k < NACC; ++k ) { acc[k] = wave_sum( acc[k] ); }
        if( lane == 0 )
        {
            if constexpr ( D )
            {
                if constexpr ( MIXED )
                {
                    // d/da = d/da' - d/d(a x) . x_i
                    acc[0] -= acc[1] * x0 + acc[2] * x1 + acc[3] * x2;
                }
                #pragma unroll
                for( int k = 0; k < 10; ++k ) { A.dx[k * A.C + i] = acc[k]; }
            }
            if constexpr ( M ) { A.diag[2 * i] = acc[10]; A.diag[2 * i + 1] = acc[11]; }
            if constexpr ( E ) { A.epart[i] = acc[12]; }
        }
    }

    // Fixed-order reduction of n partials into out[0] (one block).
    __global__ __launch_bounds__(1024) void k_reduce_sum( const double * part, int64_t n, double * out )
    {
        __shared__ double s[1024];
        const int t = threadIdx.x;
        const int64_t chunk = (n + 1023) / 1024;
        const int64_t b = (int64_t)t * chunk;
        const int64_t e = ( b + chunk < n ) ? b + chunk : n;
        double acc = 0.0;
        for( int64_t k = b; k < e; ++k ) { acc += part[k]; }
        s[t] = acc;
        __syncthreads();
        for( int off = 512; off > 0; off >>= 1 )
        {
            if( t < off ) { s[t] += s[t + off]; }
            __syncthreads();
        }
        if( t == 0 ) { out[0] = s[0]; }
    }

    // ------------------------------------------------------------------ host launchers
    template<bool E, bool D, bool M>
    static void launch_nf_flags( int kind, bool mixed, const NFArgs & A, hipStream_t s )
    {
        const dim3 grid( (unsigned)((A.L + 3) / 4) ), block( 256 );
        if( A.L == 0 ) { return; }
        if( !mixed )
        {
            if( kind == KIND_Q6P12 )      { hipLaunchKernelGGL( (k_nf_f64<KIND_Q6P12,E,D,M>), grid, block, 0, s, A ); }
            else if( kind == KIND_Q2P4 )  { hipLaunchKernelGGL( (k_nf_f64<KIND_Q2P4,E,D,M>),  grid, block, 0, s, A ); }
            else                          { hipLaunchKernelGGL( (k_nf_f64<KIND_GENERIC,E,D,M>), grid, block, 0, s, A ); }
        }
        else
        {
            if( kind == KIND_Q6P12 )      { hipLaunchKernelGGL( (k_nf_mixed<KIND_Q6P12,E,D,M>), grid, block, 0, s, A ); }
            else if( kind == KIND_Q2P4 )  { hipLaunchKernelGGL( (k_nf_mixed<KIND_Q2P4,E,D,M>),  grid, block, 0, s, A ); }
            else                          { hipLaunchKernelGGL( (k_nf_mixed<KIND_GENERIC,E,D,M>), grid, block, 0, s, A ); }
        }
    }

    template<bool E, bool D, bool M>
    static void launch_ff_flags( int kind, bool mixed, const FFArgs & A, hipStream_t s )
    {
        const dim3 grid( (unsigned)((A.C + 3) / 4) ), block( 256 );
        if( A.C == 0 ) { return; }
        if( !mixed )
        {
            if( kind == KIND_Q6P12 )      { hipLaunchKernelGGL( (k_ff<KIND_Q6P12,E,D,M,false>), grid, block, 0, s, A ); }
            else if( kind == KIND_Q2P4 )  { hipLaunchKernelGGL( (k_ff<KIND_Q2P4,E,D,M,false>),  grid, block, 0, s, A ); }
            else                          { hipLaunchKernelGGL( (k_ff<KIND_GENERIC,E,D,M,false>), grid, block, 0, s, A ); }
        }
        else
        {
            if( kind == KIND_Q6P12 )      { hipLaunchKernelGGL( (k_ff<KIND_Q6P12,E,D,M,true>), grid, block, 0, s, A ); }
            else if( kind == KIND_Q2P4 )  { hipLaunchKernelGGL( (k_ff<KIND_Q2P4,E,D,M,true>),  grid, block, 0, s, A ); }
            else                          { hipLaunchKernelGGL( (k_ff<KIND_GENERIC,E,D,M,true>), grid, block, 0, s, A ); }
        }
    }

    void launch_nf( unsigned flags, int kind, bool mixed, const NFArgs & A, hipStream_t s )
    {
        const bool E = flags & RP_ENERGY, D = flags & RP_DIFF, M = flags & RP_METRIC;
        if(  E &&  D &&  M ) { launch_nf_flags<true ,true ,true >( kind, mixed, A, s ); }
        else if(  E &&  D && !M ) { launch_nf_flags<true ,true ,false>( kind, mixed, A, s ); }
        else if(  E && !D &&  M ) { launch_nf_flags<true ,false,true >( kind, mixed, A, s ); }
        else if(  E && !D && !M ) { launch_nf_flags<true ,false,false>( kind, mixed, A, s ); }
        else if( !E &&  D &&  M ) { launch_nf_flags<false,true ,true >( kind, mixed, A, s ); }
        else if( !E &&  D && !M ) { launch_nf_flags<false,true ,false>( kind, mixed, A, s ); }
        else if( !E && !D &&  M ) { launch_nf_flags<false,false,true >( kind, mixed, A, s ); }
    }

    void launch_ff( unsigned flags, int kind, bool mixed, const FFArgs & A, hipStream_t s )
    {
        const bool E = flags & RP_ENERGY, D = flags & RP_DIFF, M = flags & RP_METRIC;
        if(  E &&  D &&  M ) { launch_ff_flags<true ,true ,true >( kind, mixed, A, s ); }
        else if(  E &&  D && !M ) { launch_ff_flags<true ,true ,false>( kind, mixed, A, s ); }
        else if(  E && !D &&  M ) { launch_ff_flags<true ,false,true >( kind, mixed, A, s ); }
        else if(  E && !D && !M ) { launch_ff_flags<true ,false,false>( kind, mixed, A, s ); }
        else if( !E &&  D &&  M ) { launch_ff_flags<false,true ,true >( kind, mixed, A, s ); }
        else if( !E &&  D && !M ) { launch_ff_flags<false,true ,false>( kind, mixed, A, s ); }
        else if( !E && !D &&  M ) { launch_ff_flags<false,false,true >( kind, mixed, A, s ); }
    }

    void launch_reduce_sum( const double * part, int64_t n, double * out, hipStream_t s )
    {
        hipLaunchKernelGGL( k_reduce_sum, dim3(1), dim3(1024), 0, s, part, n, out );
    }

} // namespace rp
