// Near-field (primitive pairs) kernels of TangentPointEnergy0 and TangentPointMetric0:
// energy, moment derivatives, metric blocks (FP64 and FP32-mixed) and the near-field block SpMM.
//
// Work unit: one workgroup (one wavefront) per row leaf I (n_I <= 32 primitives). The leaf's
// column list (the primitives of every leaf J sharing a non-admissible block with I) is processed
// in chunks of 64 columns. Each chunk is loaded cooperatively (lane l loads column l of the chunk,
// coalesced) into registers one chunk ahead and then stored into LDS, so the loads of chunk k+1
// overlap the pair arithmetic of chunk k. Lane l computes the pairs of row ii = l mod n_I of the
// leaf with the columns c, c + nS, ... of the chunk, c = l / n_I, nS = 64 / n_I. Per-lane partial
// sums are combined in LDS in a fixed order, so results are bitwise reproducible.
//
// Both orientations of every pair are visited (the row-leaf column lists cover the full symmetric
// pattern) and only row-side (S-side) derivatives are accumulated; they equal the reference's
// DX + DY sums (TP0_Kernel_NF.hpp:245-272) without scatter or atomics. The energy counts each
// unordered pair once (column > row), as the reference's upper-triangle traversal does
// (FMM_Traversor.hpp:100-101). Metric blocks are stored per leaf as [column][row] (a0,a1) pairs,
// so the lanes of a wave write consecutive slots.

#include "k_pair.hpp"
#include "rp_launch.hpp"

#include <type_traits>

namespace rp
{
    constexpr int NACC = 13;   // 0: d/da, 1..3: d/d(a hull) (scaled), 4..9: d/d(a P), 10,11: diag, 12: energy
    constexpr int CH   = 64;   // columns per chunk

    // Column data of one chunk: b, y (3), Q (6).
    template<typename R>
    struct Chunk
    {
        R       d[2][10][CH];
        int32_t j[2][CH];
        uint32_t m[2][CH];
    };

    // ------------------------------------------------------------------ energy / derivative / metric
    // MIXED: pair arithmetic in FP32 on coordinates relative to the row leaf's origin o_I (the
    // subtraction y - o_I is done in FP64 when the chunk is staged), FP32 partial sums over 8 pairs
    // flushed into FP64 accumulators. The area derivative accumulated in the local frame is moved back
    // with d/da = d/da' - sum_l d/d(a hull_l) . o_I.
    template<int KIND, bool E, bool D, bool M, bool MIXED>
    __global__ __launch_bounds__(64) void k_nf( NFArgs A )
    {
        using R = typename std::conditional<MIXED, float, double>::type;
        // FP64 metric: two more sums, the diagonal of the FP32 copies (13, 14)
        constexpr int NA = ( M && !MIXED ) ? NACC + 2 : NACC;
        union Smem { Chunk<R> ch; double red[64][NA]; };
        __shared__ Smem smem;
        Chunk<R> & sh = smem.ch;
        double (&red)[64][NA] = smem.red;
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

        double o0 = 0, o1 = 0, o2 = 0;
        if constexpr ( MIXED ) { o0 = A.origin[I]; o1 = A.origin[A.L + I]; o2 = A.origin[2 * A.L + I]; }

        // row data
        const double a_d = A.a[i];
        const R a = (R)a_d;
        R x[3], P[6];
        if constexpr ( MIXED )
        {
            x[0] = (float)(A.x[i] - o0); x[1] = (float)(A.x[T + i] - o1); x[2] = (float)(A.x[2 * T + i] - o2);
        }
        else
        {
            x[0] = A.x[i]; x[1] = A.x[T + i]; x[2] = A.x[2 * T + i];
        }
        #pragma unroll
        for( int k = 0; k < 6; ++k ) { P[k] = (R)A.P[k * T + i]; }
        const R inv_a = R(1) / a;
        const R sc = (R)A.scale;

        const int64_t cb = A.col_ptr[I];
        const int64_t M_ = A.col_ptr[I + 1] - cb;
        const int64_t vb = M ? A.val_ptr[I] : 0;

        const bool copies = M && !MIXED && A.vals32 != nullptr;
        double acc[NA];
        R part[NA];
        #pragma unroll
        for( int k = 0; k < NA; ++k ) { acc[k] = 0.0; part[k] = R(0); }
        int cnt = 0;

        // register staging of one column per lane
        R rd[10];
        int32_t rj = 0;
        uint32_t rm = 0;
        auto gload = [&]( int64_t base )
        {
            const int64_t m = base + lane;
            if( m < M_ )
            {
                rj = A.col_prim[cb + m];
                rm = A.col_mask[cb + m];
                // the column's record (a, x, P): five 16-byte loads from one place
                const double2 * q = reinterpret_cast<const double2 *>( A.rec + 10 * (int64_t)rj );
                const double2 q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3], q4 = q[4];
                rd[0] = (R)q0.x;
                if constexpr ( MIXED )
                {
                    rd[1] = (float)(q0.y - o0); rd[2] = (float)(q1.x - o1); rd[3] = (float)(q1.y - o2);
                }
                else
                {
                    rd[1] = q0.y; rd[2] = q1.x; rd[3] = q1.y;
                }
                rd[4] = (R)q2.x; rd[5] = (R)q2.y; rd[6] = (R)q3.x; rd[7] = (R)q3.y; rd[8] = (R)q4.x; rd[9] = (R)q4.y;
            }
            else
            {
                rm = 0u;
            }
        };
        auto lstore = [&]( int buf )
        {
            sh.j[buf][lane] = rj;
            sh.m[buf][lane] = rm;
            #pragma unroll
            for( int k = 0; k < 10; ++k ) { sh.d[buf][k][lane] = rd[k]; }
        };

        gload( 0 );
        lstore( 0 );
        __syncthreads();
        int buf = 0;
        for( int64_t base = 0; base < M_; base += CH )
        {
            const bool more = base + CH < M_;
            if( more ) { gload( base + CH ); }
            const int cnt_c = (int)( (M_ - base < CH) ? (M_ - base) : CH );
            if( active )
            {
                for( int mm = c; mm < cnt_c; mm += nS )
                {
                    const uint32_t mk = sh.m[buf][mm];
                    if( (mk >> ii) & 1u )
                    {
                        const int32_t j = sh.j[buf][mm];
                        const R b = sh.d[buf][0][mm];
                        const R v[3] = { sh.d[buf][1][mm] - x[0], sh.d[buf][2][mm] - x[1], sh.d[buf][3][mm] - x[2] };
                        R Q[6];
                        #pragma unroll
                        for( int k = 0; k < 6; ++k ) { Q[k] = sh.d[buf][4 + k][mm]; }
                        PairTerms<R> t;
                        pair_terms<R,KIND,E,D,M,true>( A.ex, v, P, Q, t );
                        if constexpr ( E ) { if( j > i ) { part[12] += a * t.E * b; } }
                        if constexpr ( D )
                        {
                            const R gx = t.g[0] * x[0] + t.g[1] * x[1] + t.g[2] * x[2];
                            part[0] += b * ( t.E0 + gx );
                            const R sb = sc * b;
                            part[1] -= sb * t.g[0];
                            part[2] -= sb * t.g[1];
                            part[3] -= sb * t.g[2];
                            const R bk = b * t.hK;
                            #pragma unroll
                            for( int k = 0; k < 6; ++k ) { part[4 + k] += bk * t.Vk[k]; }
                        }
                        if constexpr ( M )
                        {
                            const int64_t slot = vb + (base + mm) * nI + ii;
                            if constexpr ( MIXED ) { reinterpret_cast<float2 *>(A.vals)[slot] = make_float2( t.a0, t.a1 ); }
                            else                   { reinterpret_cast<double2 *>(A.vals)[slot] = make_double2( t.a0, t.a1 ); }
                            const R boa = b * inv_a;
                            part[10] -= boa * t.a0;
                            part[11] -= boa * t.a1;
                            if constexpr ( !MIXED )
                            {
                                if( copies )
                                {
                                    const float2 v32 = make_float2( (float)t.a0, (float)t.a1 );
                                    A.vals32[slot] = v32;
                                    part[NA - 2] -= boa * (double)v32.x;
                                    part[NA - 1] -= boa * (double)v32.y;
                                }
                            }
                        }
                        if constexpr ( MIXED )
                        {
                            if( ++cnt == 8 )
                            {
                                #pragma unroll
                                for( int k = 0; k < NACC; ++k ) { acc[k] += (double)part[k]; part[k] = 0.0f; }
                                cnt = 0;
                            }
                        }
                    }
                    else if constexpr ( M )
                    {
                        const int64_t slot = vb + (base + mm) * nI + ii;
                        if constexpr ( MIXED ) { reinterpret_cast<float2 *>(A.vals)[slot] = make_float2( 0.0f, 0.0f ); }
                        else
                        {
                            reinterpret_cast<double2 *>(A.vals)[slot] = make_double2( 0.0, 0.0 );
                            if( copies ) { A.vals32[slot] = make_float2( 0.0f, 0.0f ); }
                        }
                    }
                }
            }
            if( more ) { lstore( buf ^ 1 ); }
            __syncthreads();
            buf ^= 1;
        }
        #pragma unroll
        for( int k = 0; k < NA; ++k ) { acc[k] += (double)part[k]; }
        #pragma unroll
        for( int k = 0; k < NA; ++k ) { red[lane][k] = acc[k]; }
        __syncthreads();
        if( lane < nI )
        {
            double s[12];
            #pragma unroll
            for( int k = 0; k < 12; ++k ) { s[k] = 0.0; }
            for( int cc = 0; cc < nS; ++cc )
            {
                #pragma unroll
                for( int k = 0; k < 12; ++k ) { s[k] += red[lane + cc * nI][k]; }
            }
            if constexpr ( D )
            {
                if constexpr ( MIXED )
                {
                    // back to the global frame: every one of the (n+1) hull slots carries s[1..3]
                    s[0] -= (double)A.size * ( s[1] * o0 + s[2] * o1 + s[3] * o2 );
                }
                #pragma unroll
                for( int k = 0; k < 10; ++k ) { A.dx[k * T + i] = s[k]; }
            }
            if constexpr ( M )
            {
                A.diag[2 * i]     = s[10];
                A.diag[2 * i + 1] = s[11];
            }
            if constexpr ( M && !MIXED )
            {
                if( copies )
                {
                    double d0 = 0.0, d1 = 0.0;
                    for( int cc = 0; cc < nS; ++cc ) { d0 += red[lane + cc * nI][NA - 2]; d1 += red[lane + cc * nI][NA - 1]; }
                    A.diag32[2 * i]     = d0;
                    A.diag32[2 * i + 1] = d1;
                }
            }
        }
        if constexpr ( E )
        {
            if( lane == 0 )
            {
                double e = 0.0;
                for( int l = 0; l < 64; ++l ) { e += red[l][12]; }
                A.epart[I] = e;
            }
        }
    }

    // ------------------------------------------------------------------ near-field block SpMM
    // y_i = sum_j diag(a0, a1, a1, a1) x_j + diag_i x_i, x and y 4 x 3 per primitive (AoS, 12 doubles).
    // Epilogue (post-operator, ClusterTree::Post): ytot_i = y_i + far-field leaf total of leaf(i), then
    // S_out[i][s][c] = sum_r pre_vals[i][r][s] ytot_i[r][c] for the (n+1) vertex slots s.
    // This form (two rows per lane) serves the FP64 values (VT = double2); the FP32 copies run
    // k_nf_spmm_rows below, where one row per lane measured faster (1.47 -> 1.40 ms at 1.31M
    // triangles) while for the FP64 values it made no difference.
    template<typename VT, int SIZE>
    __global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_nf_spmm( int64_t L, const int32_t * leaf_begin, const int32_t * leaf_end,
        const int64_t * col_ptr, const int32_t * col_prim, const uint32_t * col_mask, const int64_t * val_ptr,
        const VT * vals, const double * diag, const double * xin,
        const double * leaf_out, int use_far, const double * pre_vals, double * S_out )
    {
        // 32 columns per chunk and the row sums by lane shuffles: 6.5 KB of LDS per wavefront, so that
        // 24 leaves share a CU
        constexpr int CS = 32;
        // input blocks piece by piece, sx[buffer][16-byte piece][column]: a lane reads a column's block
        // with six 16-byte loads, the columns a read instruction addresses are consecutive 16-byte
        // slots, and each gather lane writes into a contiguous run (no LDS bank conflicts)
        __shared__ double2 sx[2][6][CS];
        const int lane = threadIdx.x;
        const int64_t I = xcd_leaf( L );
        if( I < 0 ) { return; }
        const int ib = leaf_begin[I];
        const int nI = leaf_end[I] - ib;
        // a lane takes two rows of the leaf, p and p + h, for one column of the chunk at a time: every
        // column block read from LDS serves two pairs
        const int h = ( nI + 1 ) >> 1;
        const int nS = 64 / h;
        const int p = lane % h;
        const int c = lane / h;
        const bool active = lane < nS * h;
        const int r1 = p + h;
        const bool two = r1 < nI;
        const int64_t cb = col_ptr[I];
        const int64_t M_ = col_ptr[I + 1] - cb;
        const int64_t vb = val_ptr[I];
        double acc0[12], acc1[12];
        #pragma unroll
        for( int k = 0; k < 12; ++k ) { acc0[k] = 0.0; acc1[k] = 0.0; }
        // lane l loads half of the input block of column l mod 32 of the chunk (doubles 6 (l / 32) ..); the
        // column indices are fetched a chunk before the blocks they address, so the gather of the next
        // chunk waits for one load, not two
        double rx[6];
        const int half = lane >> 5, lc = lane & ( CS - 1 );
        auto iload = [&]( int64_t base ) -> int32_t
        {
            const int64_t m = base + lc;
            return ( m < M_ ) ? col_prim[cb + m] : -1;
        };
        auto gload = [&]( int32_t rj )
        {
            if( rj >= 0 )
            {
                const double2 * xj = reinterpret_cast<const double2 *>( xin + 12 * (int64_t)rj + 6 * half );
                #pragma unroll
                for( int q = 0; q < 3; ++q ) { const double2 t = xj[q]; rx[2 * q] = t.x; rx[2 * q + 1] = t.y; }
            }
        };
        auto lstore = [&]( int buf )
        {
            #pragma unroll
            for( int q = 0; q < 3; ++q ) { sx[buf][3 * half + q][lc] = make_double2( rx[2 * q], rx[2 * q + 1] ); }
        };
        int32_t rj_next = iload( CS );
        gload( iload( 0 ) );
        lstore( 0 );
        __syncthreads();
        int buf = 0;
        for( int64_t base = 0; base < M_; base += CS )
        {
            const bool more = base + CS < M_;
            if( more )
            {
                gload( rj_next );
                rj_next = iload( base + 2 * CS );
            }
            const int cnt_c = (int)( (M_ - base < CS) ? (M_ - base) : CS );
            if( active )
            {
                // the value pairs of U columns (2U slots) are loaded before the products: 64 bytes in flight
                // per lane for double2 (U = 2) and for float2 (U = 4); every tile slot exists and the
                // assembly stores 0 in the slots of masked pairs, so the products need no mask; each row
                // sums its products in column order
                constexpr int U = sizeof(VT) == 16 ? 2 : 4;
                const VT * vrow = vals + vb + base * nI + p;
                for( int mm = c; mm < cnt_c; mm += U * nS )
                {
                    VT a0[U], a1[U];   // [column u], rows p and p + h
                    #pragma unroll
                    for( int u = 0; u < U; ++u )
                    {
                        const int m2 = mm + u * nS;
                        a0[u] = VT{}; a1[u] = VT{};
                        if( m2 < cnt_c )
                        {
                            a0[u] = vrow[(int64_t)m2 * nI];
                            if( two ) { a1[u] = vrow[(int64_t)m2 * nI + h]; }
                        }
                    }
                    #pragma unroll
                    for( int u = 0; u < U; ++u )
                    {
                        const int m2 = mm + u * nS;
                        if( m2 >= cnt_c ) { break; }
                        const double b0 = (double)a0[u].x, b1 = (double)a0[u].y;
                        const double e0 = (double)a1[u].x, e1 = (double)a1[u].y;
                        #pragma unroll
                        for( int q = 0; q < 6; ++q )
                        {
                            const double2 x2 = sx[buf][q][m2];
                            const int k = 2 * q;
                            acc0[k]     += ( k < 3 ? b0 : b1 ) * x2.x;
                            acc1[k]     += ( k < 3 ? e0 : e1 ) * x2.x;
                            acc0[k + 1] += ( k + 1 < 3 ? b0 : b1 ) * x2.y;
                            acc1[k + 1] += ( k + 1 < 3 ? e0 : e1 ) * x2.y;
                        }
                    }
                }
            }
            if( more ) { lstore( buf ^ 1 ); }
            __syncthreads();
            buf ^= 1;
        }
        // row sums over the column groups of each row pair, in group order (every lane takes part in the
        // shuffles); lane q < nI then holds row q: its own first row if q < h, else the second row of q - h
        double s[12];
        #pragma unroll
        for( int k = 0; k < 12; ++k )
        {
            double s0 = 0.0, s1 = 0.0;
            for( int cc = 0; cc < nS; ++cc )
            {
                const int src = p + cc * h;
                s0 += __shfl( acc0[k], src, 64 );
                s1 += __shfl( acc1[k], src, 64 );
            }
            const double t = __shfl( s1, ( lane >= h ) ? lane - h : lane, 64 );
            s[k] = ( lane < h ) ? s0 : t;
        }
        if( lane < nI )
        {
            const int64_t i = ib + lane;
            const double d0 = diag[2 * i], d1 = diag[2 * i + 1];
            const double * xi = xin + 12 * i;
            #pragma unroll
            for( int k = 0; k < 3; ++k )  { s[k] += d0 * xi[k]; }
            #pragma unroll
            for( int k = 3; k < 12; ++k ) { s[k] += d1 * xi[k]; }
            if( use_far )
            {
                const double * lo = leaf_out + 12 * I;
                #pragma unroll
                for( int k = 0; k < 12; ++k ) { s[k] = s[k] + lo[k]; }
            }
            const double * pv = pre_vals + 4 * SIZE * i;
            #pragma unroll
            for( int sl = 0; sl < SIZE; ++sl )
            {
                #pragma unroll
                for( int cc = 0; cc < 3; ++cc )
                {
                    double t = 0.0;
                    #pragma unroll
                    for( int r = 0; r < 4; ++r ) { t += pv[SIZE * r + sl] * s[3 * r + cc]; }
                    S_out[(SIZE * i + sl) * 3 + cc] = t;
                }
            }
        }
    }

    // The same product over the FP32 copies of the values (the CG iterations of RP_MIXED_REFINED), one
    // row per lane: half the accumulators of the two-row form, so that 5 instead of 4 wavefronts share
    // a SIMD (92 registers), and 8 columns of values (64 bytes) in flight per lane. Each lane sums its
    // row's products in column order; the row sums over the column groups follow group order.
    template<int SIZE>
    __global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5))) void k_nf_spmm_rows( int64_t L, const int32_t * leaf_begin, const int32_t * leaf_end,
        const int64_t * col_ptr, const int32_t * col_prim, const int64_t * val_ptr,
        const float2 * vals, const double * diag, const double * xin,
        const double * leaf_out, int use_far, const double * pre_vals, double * S_out )
    {
        constexpr int CS = 32;
        constexpr int U = 8;
        // input blocks piece by piece, sx[buffer][16-byte piece][column]: the columns a read instruction
        // addresses are consecutive 16-byte slots and each gather lane writes into a contiguous run
        __shared__ double2 sx[2][6][CS];
        const int lane = threadIdx.x;
        const int64_t I = xcd_leaf( L );
        if( I < 0 ) { return; }
        const int ib = leaf_begin[I];
        const int nI = leaf_end[I] - ib;
        const int nS = 64 / nI;
        const int p = lane % nI;
        const int c = lane / nI;
        const bool active = lane < nS * nI;
        const int64_t cb = col_ptr[I];
        const int64_t M_ = col_ptr[I + 1] - cb;
        const int64_t vb = val_ptr[I];
        double acc[12];
        #pragma unroll
        for( int k = 0; k < 12; ++k ) { acc[k] = 0.0; }
        // input-block staging as in k_nf_spmm: lane l gathers half of the block of column l mod 32, the
        // column indices a chunk ahead
        double rx[6];
        const int half = lane >> 5, lc = lane & ( CS - 1 );
        auto iload = [&]( int64_t base ) -> int32_t
        {
            const int64_t m = base + lc;
            return ( m < M_ ) ? col_prim[cb + m] : -1;
        };
        auto gload = [&]( int32_t rj )
        {
            if( rj >= 0 )
            {
                const double2 * xj = reinterpret_cast<const double2 *>( xin + 12 * (int64_t)rj + 6 * half );
                #pragma unroll
                for( int q = 0; q < 3; ++q ) { const double2 t = xj[q]; rx[2 * q] = t.x; rx[2 * q + 1] = t.y; }
            }
        };
        auto lstore = [&]( int buf )
        {
            #pragma unroll
            for( int q = 0; q < 3; ++q ) { sx[buf][3 * half + q][lc] = make_double2( rx[2 * q], rx[2 * q + 1] ); }
        };
        int32_t rj_next = iload( CS );
        gload( iload( 0 ) );
        lstore( 0 );
        __syncthreads();
        int buf = 0;
        for( int64_t base = 0; base < M_; base += CS )
        {
            const bool more = base + CS < M_;
            if( more )
            {
                gload( rj_next );
                rj_next = iload( base + 2 * CS );
            }
            const int cnt_c = (int)( (M_ - base < CS) ? (M_ - base) : CS );
            if( active )
            {
                // the chunk's values from one wave-uniform base with 32-bit lane offsets (a tile holds at
                // most 32 rows x its column count of float2 values, far below 2^31 bytes)
                const char * vchunk = reinterpret_cast<const char *>( vals + vb + base * nI );
                const uint32_t ostep = (uint32_t)( nS * nI ) * (uint32_t)sizeof(float2);
                for( int mm = c; mm < cnt_c; mm += U * nS )
                {
                    float2 a[U];
                    uint32_t ob = (uint32_t)( mm * nI + p ) * (uint32_t)sizeof(float2);
                    #pragma unroll
                    for( int u = 0; u < U; ++u )
                    {
                        // a slot past the chunk's last column is never multiplied, so it needs no value
                        if( mm + u * nS < cnt_c ) { a[u] = *reinterpret_cast<const float2 *>( vchunk + ob ); }
                        ob += ostep;
                    }
                    #pragma unroll
                    for( int u = 0; u < U; ++u )
                    {
                        const int m2 = mm + u * nS;
                        if( m2 >= cnt_c ) { break; }
                        const double b0 = (double)a[u].x, b1 = (double)a[u].y;
                        #pragma unroll
                        for( int q = 0; q < 6; ++q )
                        {
                            const double2 x2 = sx[buf][q][m2];
                            const int k = 2 * q;
                            acc[k]     += ( k < 3 ? b0 : b1 ) * x2.x;
                            acc[k + 1] += ( k + 1 < 3 ? b0 : b1 ) * x2.y;
                        }
                    }
                }
            }
            if( more ) { lstore( buf ^ 1 ); }
            __syncthreads();
            buf ^= 1;
        }
        // row sums over the column groups in group order (every lane takes part); lane q < nI holds row q
        double s[12];
        #pragma unroll
        for( int k = 0; k < 12; ++k )
        {
            double s0 = 0.0;
            for( int cc = 0; cc < nS; ++cc ) { s0 += __shfl( acc[k], p + cc * nI, 64 ); }
            s[k] = s0;
        }
        if( lane < nI )
        {
            const int64_t i = ib + lane;
            const double d0 = diag[2 * i], d1 = diag[2 * i + 1];
            const double * xi = xin + 12 * i;
            #pragma unroll
            for( int k = 0; k < 3; ++k )  { s[k] += d0 * xi[k]; }
            #pragma unroll
            for( int k = 3; k < 12; ++k ) { s[k] += d1 * xi[k]; }
            if( use_far )
            {
                const double * lo = leaf_out + 12 * I;
                #pragma unroll
                for( int k = 0; k < 12; ++k ) { s[k] = s[k] + lo[k]; }
            }
            const double * pv = pre_vals + 4 * SIZE * i;
            #pragma unroll
            for( int sl = 0; sl < SIZE; ++sl )
            {
                #pragma unroll
                for( int cc = 0; cc < 3; ++cc )
                {
                    double t = 0.0;
                    #pragma unroll
                    for( int r = 0; r < 4; ++r ) { t += pv[SIZE * r + sl] * s[3 * r + cc]; }
                    S_out[(SIZE * i + sl) * 3 + cc] = t;
                }
            }
        }
    }

    // ------------------------------------------------------------------ host launchers
    template<bool E, bool D, bool M>
    static void launch_nf_flags( int kind, bool mixed, const NFArgs & A, hipStream_t s )
    {
        if( A.L == 0 ) { return; }
        const dim3 grid = leaf_grid( A.L ), block( 64 );
        if( !mixed )
        {
            if( kind == KIND_Q6P12 )      { hipLaunchKernelGGL( (k_nf<KIND_Q6P12,E,D,M,false>),  grid, block, 0, s, A ); }
            else if( kind == KIND_Q2P4 )  { hipLaunchKernelGGL( (k_nf<KIND_Q2P4,E,D,M,false>),   grid, block, 0, s, A ); }
            else                          { hipLaunchKernelGGL( (k_nf<KIND_GENERIC,E,D,M,false>), grid, block, 0, s, A ); }
        }
        else
        {
            if( kind == KIND_Q6P12 )      { hipLaunchKernelGGL( (k_nf<KIND_Q6P12,E,D,M,true>),  grid, block, 0, s, A ); }
            else if( kind == KIND_Q2P4 )  { hipLaunchKernelGGL( (k_nf<KIND_Q2P4,E,D,M,true>),   grid, block, 0, s, A ); }
            else                          { hipLaunchKernelGGL( (k_nf<KIND_GENERIC,E,D,M,true>), grid, block, 0, s, A ); }
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

    void launch_nf_spmm_post( bool mixed, int size, int64_t T, int64_t L, const int32_t * lb, const int32_t * le,
                              const int64_t * cp, const int32_t * cpr, const uint32_t * cm, const int64_t * vp,
                              const void * vals, const double * diag, const double * xin,
                              const double * leaf_out, int use_far, const double * pre_vals, double * S_out, hipStream_t s )
    {
        if( !L ) { return; }
        const dim3 grid = leaf_grid( L ), block( 64 );
        if( size == 3 )
        {
            if( mixed ) { hipLaunchKernelGGL( (k_nf_spmm_rows<3>), grid, block, 0, s, L, lb, le, cp, cpr, vp, (const float2 *)vals, diag, xin, leaf_out, use_far, pre_vals, S_out ); }
            else        { hipLaunchKernelGGL( (k_nf_spmm<double2,3>), grid, block, 0, s, L, lb, le, cp, cpr, cm, vp, (const double2 *)vals, diag, xin, leaf_out, use_far, pre_vals, S_out ); }
        }
        else
        {
            if( mixed ) { hipLaunchKernelGGL( (k_nf_spmm_rows<2>), grid, block, 0, s, L, lb, le, cp, cpr, vp, (const float2 *)vals, diag, xin, leaf_out, use_far, pre_vals, S_out ); }
            else        { hipLaunchKernelGGL( (k_nf_spmm<double2,2>), grid, block, 0, s, L, lb, le, cp, cpr, cm, vp, (const double2 *)vals, diag, xin, leaf_out, use_far, pre_vals, S_out ); }
        }
    }

} // namespace rp
