// Far-field (cluster pairs) kernel of TangentPointEnergy0: energy, moment derivatives and metric
// blocks, FP64 and FP32-mixed; fixed-order energy reduction.
//
// One wavefront per cluster row of the full symmetric far pattern; lanes stride the row's columns;
// deterministic shuffle reduction. Like the near-field kernel it traverses both orientations of every
// pair and accumulates only row-side (S-side) derivatives, which equal the reference's DX + DY sums
// (TP0_Kernel_FF.hpp:234-254) without scatter or atomics. The energy counts each unordered pair once
// (column > row), as the reference's upper-triangle traversal does (FMM_Traversor.hpp:100-101).

#include "k_pair.hpp"
#include "rp_launch.hpp"

namespace rp
{
    constexpr int NACC = 13;   // 0: d/da, 1..3: d/d(a x) (scaled), 4..9: d/d(a P), 10,11: diag, 12: energy

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

    // Fixed-order two-level reduction: block b sums the contiguous range [b*chunk, (b+1)*chunk) of the
    // partials into out[1 + b] (tree in LDS), then one block sums out[1 .. nb] into out[0].
    __global__ __launch_bounds__(256) void k_reduce_blocks( const double * part, int64_t n, int64_t chunk, double * out )
    {
        __shared__ double s[256];
        const int t = threadIdx.x;
        const int64_t b0 = (int64_t)blockIdx.x * chunk;
        const int64_t e0 = ( b0 + chunk < n ) ? b0 + chunk : n;
        double acc = 0.0;
        for( int64_t k = b0 + t; k < e0; k += 256 ) { acc += part[k]; }
        s[t] = acc;
        __syncthreads();
        for( int off = 128; off > 0; off >>= 1 )
        {
            if( t < off ) { s[t] += s[t + off]; }
            __syncthreads();
        }
        if( t == 0 ) { out[1 + blockIdx.x] = s[0]; }
    }

    __global__ __launch_bounds__(256) void k_reduce_final( int64_t nb, double * out )
    {
        __shared__ double s[256];
        const int t = threadIdx.x;
        double acc = 0.0;
        for( int64_t k = t; k < nb; k += 256 ) { acc += out[1 + k]; }
        s[t] = acc;
        __syncthreads();
        for( int off = 128; off > 0; off >>= 1 )
        {
            if( t < off ) { s[t] += s[t + off]; }
            __syncthreads();
        }
        if( t == 0 ) { out[0] = s[0]; }
    }

    // ------------------------------------------------------------------ host launchers
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
        // out must hold 1 + REDUCE_BLOCKS doubles
        const int64_t nb = std::min<int64_t>( REDUCE_BLOCKS, std::max<int64_t>( 1, (n + 4095) / 4096 ) );
        const int64_t chunk = (n + nb - 1) / nb;
        hipLaunchKernelGGL( k_reduce_blocks, dim3((unsigned)nb), dim3(256), 0, s, part, n, chunk, out );
        hipLaunchKernelGGL( k_reduce_final, dim3(1), dim3(256), 0, s, nb, out );
    }

} // namespace rp
