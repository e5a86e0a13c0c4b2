// Triangular solves with the supernodal Cholesky factor of the H^1 metric, H = L L^T, for three
// right-hand sides (row-major n x 3). The reference solves with its sparse Cholesky factor on the host
// (src/SimplicialMesh/H1Metric.hpp:208-270, called twice per preconditioner application from
// src/Metrics/TangentPointMetric0.hpp:88-126).
//
// Factor layout on the device (one column-major r x w block per supernode s, rows R_s sorted, the first
// w of them the supernode's own columns f..f+w-1): the top w x w block holds inv(L_ss) (lower
// triangle), the rows below hold M = L_off inv(L_ss), formed once per factor by k_chol_form_m from the
// uploaded L_off. With it both sweeps need one pass per elimination level:
//
//   forward (leaves to root):  z_s starts as [b_s; 0] plus the children's update vectors (in child
//       order; each front row pulls the child update rows that map to it, from lists built at upload), then
//           y_s = inv(L_ss) z_s[0:w],   z_s[w:r] -= M z_s[0:w]      (= L_off y_s)
//       z_s[0:w] keeps y_s for the backward sweep;
//   backward (root to leaves): x_s = inv(L_ss)^T y_s - M^T x(R_s[w:]), i.e. for column c the dot product
//       of the contiguous block column c (rows c..r-1) with v = [y_s; -x(R_s[w:])].
//
// Levels whose supernodes have at most 64 rows run a group of 16, 32 or 64 lanes per supernode (a lane
// per front row forward, a lane per row of v and a group sum per column backward); larger levels run workgroups
// over tasks (supernode, block of 64 front rows) forward, 4 or 16 waves splitting the columns, and
// (supernode, group of columns) backward, a group of lanes per column. Every sum has a fixed order, so
// the solve is deterministic and uses no atomics.

#include "rp_launch.hpp"

namespace rp
{
    constexpr int CHOL_QC = 512;   // rows of v staged in LDS per chunk of the backward tasks (rows beyond CHOL_BWD_ONE)
    constexpr int CHOL_BWD_ONE = 2560;   // backward tasks of supernodes with at most this many rows stage v at once
    // block values a lane loads ahead of the front merge (forward) or the staging of v (backward) on the
    // wide levels (16 against 32: 15 us less per preconditioner application, the same sums; 8 ties, 24 slower)
    constexpr int CHOL_PF = 16;

    __device__ __forceinline__ int chol_pow2_at_least( int x )
    {
        int g = 1;
        while( g < x && g < 64 ) { g <<= 1; }
        return g;
    }

    // Sum of v over aligned groups of G lanes (G a power of two <= 64); every lane takes part.
    __device__ __forceinline__ double group_sum( double v, int G )
    {
        for( int off = G >> 1; off > 0; off >>= 1 ) { v += __shfl_xor( v, off, 64 ); }
        return v;
    }

    // ---------------------------------------------------------------------------------------------
    // M = L_off inv(L_ss) per supernode (tasks: supernode, first update row); the top block is copied.
    // One lane per update row, columns split over the 4 waves.
    __global__ __launch_bounds__(256) void k_chol_form_m( const CholTask * tasks, const double * Lin, double * Lout )
    {
        const CholTask tk = tasks[blockIdx.x];
        const int w = tk.w, r = tk.r;
        const int64_t vb = tk.vp;
        const double * Li = Lin + vb;
        double * Lo = Lout + vb;
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        if( tk.k == 0 )
        {
            // copy inv(L_ss) (the whole top w x w block, column by column)
            for( int64_t e = threadIdx.x; e < (int64_t)w * w; e += 256 )
            {
                const int c = (int)( e / w ), a = (int)( e - (int64_t)c * w );
                Lo[(int64_t)c * r + a] = Li[(int64_t)c * r + a];
            }
        }
        const int q = w + tk.k + lane;   // front row
        if( q >= r ) { return; }
        const int cw = ( w + 3 ) / 4;
        const int cb = wv * cw, ce = min( w, cb + cw );
        for( int c = cb; c < ce; ++c )
        {
            // M(q, c) = sum_{a >= c} L_off(q, a) inv(a, c)
            double acc = 0.0;
            const double * invc = Li + (int64_t)c * r;
            for( int a = c; a < w; ++a ) { acc += Li[(int64_t)a * r + q] * invc[a]; }
            Lo[(int64_t)c * r + q] = acc;
        }
    }

    // ---------------------------------------------------------------------------------------------
    // The merged front value of row a of supernode s: the supernode's own right-hand side (rows a < w,
    // already multiplied by alpha and permuted) or 0, plus the children's update rows that map to it, in
    // child order (pull entries built at upload, the same additions in the same order as pushing the
    // children one after another).
    __device__ __forceinline__ void chol_front_row( const CholDev & F, int64_t rb, int f, int w, int a, double v[3] )
    {
        const int2 pr = F.pull2[rb + a];
        if( a < w ) { for( int k = 0; k < 3; ++k ) { v[k] = F.bp[3 * (int64_t)( f + a ) + k]; } }
        else        { for( int k = 0; k < 3; ++k ) { v[k] = 0.0; } }
        if( pr.x >= 0 )
        {
            double u0[3], u1[3] = { 0.0, 0.0, 0.0 };
            for( int k = 0; k < 3; ++k ) { u0[k] = F.z[3 * (int64_t)pr.x + k]; }
            if( pr.y >= 0 ) { for( int k = 0; k < 3; ++k ) { u1[k] = F.z[3 * (int64_t)pr.y + k]; } }
            for( int k = 0; k < 3; ++k ) { v[k] += u0[k]; }
            if( pr.y >= 0 ) { for( int k = 0; k < 3; ++k ) { v[k] += u1[k]; } }
        }
        else if( pr.x < -1 )
        {
            for( int32_t e = -2 - pr.x; e < pr.y; ++e )
            {
                const int64_t q = F.pull_src[e];
                for( int k = 0; k < 3; ++k ) { v[k] += F.z[3 * q + k]; }
            }
        }
    }

    // Bp = alpha B(perm) in the factor's numbering.
    __global__ __launch_bounds__(256) void k_chol_gather( int64_t n, const int32_t * perm, double alpha, const double * B, int64_t ldB, double * bp )
    {
        const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
        if( j >= n ) { return; }
        const int64_t o = perm[j];
        for( int k = 0; k < 3; ++k ) { bp[3 * j + k] = alpha * B[ldB * o + k]; }
    }

    // Forward, small levels (at most CHOL_SMALL_ROWS rows per supernode): a group of G lanes per
    // supernode (64 / G supernodes per wavefront), lane a taking the front rows a, a + G, ...; G follows
    // the level's typical front, not its largest, so that few lanes idle.
    template<int G>
    __global__ __launch_bounds__(256) void k_chol_fwd_small( CholDev F, const CholTask * desc, int64_t count, int rs )
    {
        extern __shared__ double fr_all[];   // (256 / G) x 3 rs, rs = the level's largest row count
        const int lane = threadIdx.x % G;
        const int64_t t = ( (int64_t)blockIdx.x * 256 + threadIdx.x ) / G;
        if( t >= count ) { return; }
        double * fr = fr_all + 3 * rs * (int)( threadIdx.x / G );
        const CholTask d = desc[t];
        const int f = d.f, w = d.w, r = d.r;
        const int64_t rb = d.rb;
        const double * Ls = F.vals + d.vp;
        // the first row's first block values go out before the front merge
        constexpr int PFS = 16;
        const int ce0 = ( lane < r ) ? ( ( lane < w ) ? lane + 1 : w ) : 0;
        double lp[PFS];
        #pragma unroll
        for( int u = 0; u < PFS; ++u ) { lp[u] = ( u < ce0 ) ? Ls[(int64_t)u * r + lane] : 0.0; }
        for( int a = lane; a < r; a += G )
        {
            double v[3];
            chol_front_row( F, rb, f, w, a, v );
            for( int k = 0; k < 3; ++k ) { fr[3 * a + k] = v[k]; }
        }
        __builtin_amdgcn_fence( __ATOMIC_RELEASE, "wavefront" );
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "wavefront" );
        double * z = F.z + 3 * rb;
        for( int a = lane; a < r; a += G )
        {
            const int ce = ( a < w ) ? a + 1 : w;
            double acc[3] = { 0.0, 0.0, 0.0 };
            int c = 0;
            if( a == lane )
            {
                #pragma unroll
                for( int u = 0; u < PFS; ++u )
                {
                    if( u < ce ) { for( int k = 0; k < 3; ++k ) { acc[k] += lp[u] * fr[3 * u + k]; } }
                }
                c = PFS;
            }
            // the remaining block values of the row in guarded batches of eight loads (sums in column order)
            for( ; c < ce; c += 8 )
            {
                double lv[8];
                #pragma unroll
                for( int u = 0; u < 8; ++u ) { lv[u] = ( c + u < ce ) ? Ls[(int64_t)( c + u ) * r + a] : 0.0; }
                #pragma unroll
                for( int u = 0; u < 8; ++u )
                {
                    if( c + u < ce ) { for( int k = 0; k < 3; ++k ) { acc[k] += lv[u] * fr[3 * ( c + u ) + k]; } }
                }
            }
            if( a < w ) { for( int k = 0; k < 3; ++k ) { z[3 * a + k] = acc[k]; } }
            else        { for( int k = 0; k < 3; ++k ) { z[3 * a + k] = fr[3 * a + k] - acc[k]; } }
        }
    }

    // Forward, larger levels: one workgroup of NW waves per task (supernode, 64 front rows); the waves
    // split the columns and add their partial sums in LDS. NW = 1, 2 or 4 serves levels whose supernodes
    // are at most CHOL_NARROW_WIDTH columns wide, 16 the wider ones (the launch picks NW by the level's
    // widest supernode).
    template<int NW>
    __global__ __launch_bounds__(64 * NW) void k_chol_fwd_big( CholDev F, const CholTask * tasks )
    {
        constexpr int NT = 64 * NW;
        __shared__ double zt[3 * ( NW == 16 ? CHOL_MAX_WIDTH : CHOL_NARROW_WIDTH )];
        __shared__ double part[NW * 64 * 3];
        const CholTask tk = tasks[blockIdx.x];
        const int row0 = tk.k, f = tk.f, w = tk.w, r = tk.r;
        const int64_t rb = tk.rb;
        const int tid = threadIdx.x;
        const int lane = tid & 63, wv = tid >> 6;
        const int a = row0 + lane;
        // the lane's block values do not depend on the front: their loads go out before the merge
        const double * Ls = F.vals + tk.vp;
        const int cw = ( w + NW - 1 ) / NW;
        const int cb = wv * cw;
        const int ce = ( a < r ) ? min( ( a < w ) ? a + 1 : w, cb + cw ) : cb;
        constexpr int PF = ( NW == 16 ) ? CHOL_PF : 8;   // deeper prefetch on the wide (latency-bound) levels
        double lp[PF];
        #pragma unroll
        for( int u = 0; u < PF; ++u ) { lp[u] = ( cb + u < ce ) ? Ls[(int64_t)( cb + u ) * r + a] : 0.0; }
        for( int a2 = tid; a2 < w; a2 += NT )
        {
            double v[3];
            chol_front_row( F, rb, f, w, a2, v );
            for( int k = 0; k < 3; ++k ) { zt[3 * a2 + k] = v[k]; }
        }
        double vb[3] = { 0.0, 0.0, 0.0 };
        if( wv == 0 && a >= w && a < r ) { chol_front_row( F, rb, f, w, a, vb ); }
        __syncthreads();
        double acc[3] = { 0.0, 0.0, 0.0 };
        #pragma unroll
        for( int u = 0; u < PF; ++u )
        {
            if( cb + u < ce ) { for( int k = 0; k < 3; ++k ) { acc[k] += lp[u] * zt[3 * ( cb + u ) + k]; } }
        }
        // the remaining block values in guarded batches of eight loads (sums in column order)
        for( int c = cb + PF; c < ce; c += 8 )
        {
            double lv[8];
            #pragma unroll
            for( int u = 0; u < 8; ++u ) { lv[u] = ( c + u < ce ) ? Ls[(int64_t)( c + u ) * r + a] : 0.0; }
            #pragma unroll
            for( int u = 0; u < 8; ++u )
            {
                if( c + u < ce ) { for( int k = 0; k < 3; ++k ) { acc[k] += lv[u] * zt[3 * ( c + u ) + k]; } }
            }
        }
        for( int k = 0; k < 3; ++k ) { part[( wv * 64 + lane ) * 3 + k] = acc[k]; }
        __syncthreads();
        if( wv == 0 && a < r )
        {
            double * z = F.z + 3 * rb;
            for( int k = 0; k < 3; ++k )
            {
                double sum = part[lane * 3 + k];
                #pragma unroll
                for( int u = 1; u < NW; ++u ) { sum += part[( u * 64 + lane ) * 3 + k]; }
                z[3 * a + k] = ( a < w ) ? sum : vb[k] - sum;
            }
        }
    }

    // Forward, wide levels of few supernodes, in two launches. First the merged front of every row of
    // the level's supernodes into fr (tasks: supernode, 64 rows; one lane per row), then the products
    // over tasks of 16 rows: a row's columns are split over 64 groups of at most 32 columns each (one
    // batch of block-value loads per lane), the own rows of the front staged in LDS from fr, and the 64
    // partial sums of a row added in group order.
    __global__ __launch_bounds__(64) void k_chol_fwd_front( CholDev F, const CholTask * tasks, double * fr )
    {
        const CholTask tk = tasks[blockIdx.x];
        const int a = tk.k + (int)threadIdx.x;
        if( a >= tk.r ) { return; }
        double v[3];
        chol_front_row( F, tk.rb, tk.f, tk.w, a, v );
        for( int k = 0; k < 3; ++k ) { fr[3 * ( tk.rb + a ) + k] = v[k]; }
    }

    __global__ __launch_bounds__(1024) void k_chol_fwd_split( CholDev F, const CholTask * tasks, const double * fr )
    {
        constexpr int RB = 16;          // rows per task
        constexpr int NG = 1024 / RB;   // column groups per row
        constexpr int CW = ( CHOL_MAX_WIDTH + NG - 1 ) / NG;   // columns per group at most
        __shared__ double zt[3 * CHOL_MAX_WIDTH];
        __shared__ double part[NG][RB][3];
        const CholTask tk = tasks[blockIdx.x];
        const int w = tk.w, r = tk.r;
        const int64_t rb = tk.rb;
        const int tid = threadIdx.x;
        const int i = tid % RB, g = tid / RB;
        const int a = tk.k + i;
        const double * Ls = F.vals + tk.vp;
        const int cw = ( w + NG - 1 ) / NG;
        const int cb = g * cw;
        const int ce = ( a < r ) ? min( ( a < w ) ? a + 1 : w, cb + cw ) : cb;
        // the lane's block values do not depend on the front: they go out before the staging
        double lp[CW];
        #pragma unroll
        for( int u = 0; u < CW; ++u ) { lp[u] = ( cb + u < ce ) ? Ls[(int64_t)( cb + u ) * r + a] : 0.0; }
        for( int c = tid; c < w; c += 1024 )
        {
            for( int k = 0; k < 3; ++k ) { zt[3 * c + k] = fr[3 * ( rb + c ) + k]; }
        }
        __syncthreads();
        double acc[3] = { 0.0, 0.0, 0.0 };
        #pragma unroll
        for( int u = 0; u < CW; ++u )
        {
            if( cb + u < ce ) { for( int k = 0; k < 3; ++k ) { acc[k] += lp[u] * zt[3 * ( cb + u ) + k]; } }
        }
        for( int k = 0; k < 3; ++k ) { part[g][i][k] = acc[k]; }
        __syncthreads();
        if( tid < RB && a < r )
        {
            double * z = F.z + 3 * rb;
            for( int k = 0; k < 3; ++k )
            {
                double sum = part[0][tid][k];
                for( int u = 1; u < NG; ++u ) { sum += part[u][tid][k]; }
                z[3 * a + k] = ( a < w ) ? sum : fr[3 * ( rb + a ) + k] - sum;
            }
        }
    }

    // ---------------------------------------------------------------------------------------------
    // Backward, small levels: a group of G lanes per supernode (G follows the level's mean row count);
    // the lanes stage v in LDS (rows lane, lane + G, ...), then x_c = sum over rows i >= c of L(i, c) v_i
    // for every column, each by Lc lanes (as many as the group has to spare, up to 16), reading the
    // transposed copy of the block (row-major), the first rows loaded before the staging.
    template<int G>
    __global__ __launch_bounds__(256) void k_chol_bwd_small( CholDev F, const CholTask * desc, int64_t count, int rs )
    {
        extern __shared__ double v_all[];   // (256 / G) x 3 rs, rs = the level's largest row count
        const int lane = threadIdx.x % G;
        const int64_t t = ( (int64_t)blockIdx.x * 256 + threadIdx.x ) / G;
        if( t >= count ) { return; }
        double * vs = v_all + 3 * rs * (int)( threadIdx.x / G );
        const CholTask d = desc[t];
        const int f = d.f, w = d.w, r = d.r;
        const int64_t rb = d.rb;
        const double * Lt = F.vals_t + d.vp;   // Lt[i * w + c] = L(i, c)
        // Lc lanes per column (the largest power of two up to 16 with Lc w <= G): lane (cl, sub) takes the
        // rows c + sub, c + sub + Lc, ... of the columns c = cl, cl + G / Lc, ...; the Lc partial sums of a
        // column are added by an aligned shuffle tree
        int Lc = 1;
        while( Lc < 16 && 2 * Lc * w <= G ) { Lc <<= 1; }
        const int cl = lane / Lc, sub = lane % Lc, ncs = G / Lc;
        constexpr int PFS = 16;
        double lp[PFS];
        #pragma unroll
        for( int u = 0; u < PFS; ++u )
        {
            const int i = cl + sub + u * Lc;
            lp[u] = ( cl < w && i < r ) ? Lt[(int64_t)i * w + cl] : 0.0;
        }
        for( int i = lane; i < r; i += G )
        {
            if( i < w ) { for( int k = 0; k < 3; ++k ) { vs[3 * i + k] = F.z[3 * ( rb + i ) + k]; } }
            else
            {
                const int64_t j = F.rows[rb + i];
                for( int k = 0; k < 3; ++k ) { vs[3 * i + k] = -F.xp[3 * j + k]; }
            }
        }
        __builtin_amdgcn_fence( __ATOMIC_RELEASE, "wavefront" );
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "wavefront" );
        for( int c = cl; c < w; c += ncs )
        {
            double acc[3] = { 0.0, 0.0, 0.0 };
            int i = c + sub;
            if( c == cl )
            {
                // the slot's first column: its first PFS rows of this lane were loaded before the staging
                #pragma unroll
                for( int u = 0; u < PFS; ++u )
                {
                    const int iu = c + sub + u * Lc;
                    if( iu < r ) { for( int k = 0; k < 3; ++k ) { acc[k] += lp[u] * vs[3 * iu + k]; } }
                }
                i = c + sub + PFS * Lc;
            }
            // the remaining rows of the lane in guarded batches of eight loads (sums in row order)
            for( ; i < r; i += 8 * Lc )
            {
                double lv[8];
                #pragma unroll
                for( int u = 0; u < 8; ++u ) { lv[u] = ( i + u * Lc < r ) ? Lt[(int64_t)( i + u * Lc ) * w + c] : 0.0; }
                #pragma unroll
                for( int u = 0; u < 8; ++u )
                {
                    if( i + u * Lc < r ) { for( int k = 0; k < 3; ++k ) { acc[k] += lv[u] * vs[3 * ( i + u * Lc ) + k]; } }
                }
            }
            for( int k = 0; k < 3; ++k ) { acc[k] = group_sum( acc[k], Lc ); }
            if( sub == 0 ) { for( int k = 0; k < 3; ++k ) { F.xp[3 * (int64_t)( f + c ) + k] = acc[k]; } }
        }
    }

    // Transposed (row-major) copy of the blocks of the small levels: Lt[i * w + c] = L(i, c).
    __global__ __launch_bounds__(256) void k_chol_transpose( const CholTask * desc, int64_t count, const double * vals, double * vals_t )
    {
        const int64_t t = blockIdx.x;
        if( t >= count ) { return; }
        const CholTask d = desc[t];
        const int w = d.w, r = d.r;
        for( int e = threadIdx.x; e < w * r; e += 256 )
        {
            const int i = e / w, c = e - i * w;
            vals_t[d.vp + e] = vals[d.vp + (int64_t)c * r + i];
        }
    }

    // Backward, larger levels: one workgroup per task (supernode, group of columns), a group of G lanes
    // per column. On the wide levels (few tasks, latency-bound) a task whose rows of v (r - k) fit the
    // lds_rows rows of LDS the launch allocated stages all of v at once and loads the lane's block values
    // before staging; otherwise v is staged in chunks of CHOL_QC rows (lds_rows >= CHOL_QC then). (Loading
    // the first batch of 8 before the staging also on the narrow levels measured slower: 1.787 against
    // 1.776 ms per preconditioner application at 1.31M triangles.)
    template<bool WIDE>
    __global__ __launch_bounds__(256) void k_chol_bwd_big( CholDev F, const CholTask * tasks, int lds_rows )
    {
        extern __shared__ double v[];
        const CholTask tk = tasks[blockIdx.x];
        const int f = tk.f, w = tk.w, r = tk.r;
        const int64_t rb = tk.rb;
        const double * z = F.z + 3 * rb;
        const int32_t * R = F.rows + rb;
        const int G = chol_bwd_group( r, WIDE );
        const int grp = threadIdx.x / G, gl = threadIdx.x % G;
        const int c = tk.k + grp;
        const double * col = F.vals + tk.vp + (int64_t)c * r;
        double acc[3] = { 0.0, 0.0, 0.0 };
        if( WIDE && r - tk.k <= lds_rows )
        {
            const int i0 = tk.k;          // no column of the task needs rows above its first column
            const int ib = ( c < w ) ? c : r;
            double lp[CHOL_PF];
            #pragma unroll
            for( int u = 0; u < CHOL_PF; ++u )
            {
                const int i = ib + gl + u * G;
                lp[u] = ( i < r ) ? col[i] : 0.0;
            }
            for( int i = i0 + threadIdx.x; i < r; i += 256 )
            {
                const int q = i - i0;
                if( i < w ) { for( int k = 0; k < 3; ++k ) { v[3 * q + k] = z[3 * i + k]; } }
                else
                {
                    const int64_t j = R[i];
                    for( int k = 0; k < 3; ++k ) { v[3 * q + k] = -F.xp[3 * j + k]; }
                }
            }
            __syncthreads();
            #pragma unroll
            for( int u = 0; u < CHOL_PF; ++u )
            {
                const int i = ib + gl + u * G;
                if( i < r ) { for( int k = 0; k < 3; ++k ) { acc[k] += lp[u] * v[3 * ( i - i0 ) + k]; } }
            }
            for( int i = ib + gl + CHOL_PF * G; i < r; i += 8 * G )
            {
                double lv[8];
                #pragma unroll
                for( int u = 0; u < 8; ++u ) { lv[u] = ( i + u * G < r ) ? col[i + u * G] : 0.0; }
                #pragma unroll
                for( int u = 0; u < 8; ++u )
                {
                    if( i + u * G < r ) { for( int k = 0; k < 3; ++k ) { acc[k] += lv[u] * v[3 * ( i + u * G - i0 ) + k]; } }
                }
            }
        }
        else
        {
            for( int i0 = tk.k; i0 < r; i0 += CHOL_QC )
            {
                const int in = min( CHOL_QC, r - i0 );
                __syncthreads();
                for( int i = threadIdx.x; i < in; i += 256 )
                {
                    const int ii = i0 + i;
                    if( ii < w ) { for( int k = 0; k < 3; ++k ) { v[3 * i + k] = z[3 * ii + k]; } }
                    else
                    {
                        const int64_t j = R[ii];
                        for( int k = 0; k < 3; ++k ) { v[3 * i + k] = -F.xp[3 * j + k]; }
                    }
                }
                __syncthreads();
                if( c < w )
                {
                    const int ib = max( c, i0 );
                    const int ie = i0 + in;
                    for( int i = ib + gl; i < ie; i += 8 * G )
                    {
                        double lv[8];
                        #pragma unroll
                        for( int u = 0; u < 8; ++u ) { lv[u] = ( i + u * G < ie ) ? col[i + u * G] : 0.0; }
                        #pragma unroll
                        for( int u = 0; u < 8; ++u )
                        {
                            if( i + u * G < ie ) { for( int k = 0; k < 3; ++k ) { acc[k] += lv[u] * v[3 * ( i + u * G - i0 ) + k]; } }
                        }
                    }
                }
            }
        }
        for( int k = 0; k < 3; ++k ) { acc[k] = group_sum( acc[k], G ); }
        if( c < w && gl == 0 ) { for( int k = 0; k < 3; ++k ) { F.xp[3 * (int64_t)( f + c ) + k] = acc[k]; } }
    }

    // The learned preconditioner's ends fused into the permutations (the same products as scaling
    // separately): in, bp = sqrt_d B(perm); out, with t = xp + (c / dA) sqrt_d X (the diagonal term, c = 0:
    // none), Y(perm) = alpha sqrt_d t + beta Y(perm).
    __global__ void k_chol_gather_scaled( int64_t n, const int32_t * perm, const double * sd, const double * B, int64_t ldB, double * bp )
    {
        const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
        if( j >= n ) { return; }
        const int64_t o = perm[j];
        const double c = 1.0 * sd[o];
        for( int k = 0; k < 3; ++k ) { bp[3 * j + k] = c * B[ldB * o + k]; }
    }

    __global__ void k_chol_scatter_scaled( int64_t n, const int32_t * perm, const double * xp, const double * sd, const double * X, int64_t ldX,
                                           double cw, const double * dA, double alpha, double beta, double * Y, int64_t ldY )
    {
        const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
        if( j >= n ) { return; }
        const int64_t o = perm[j];
        const double s1 = 1.0 * sd[o];
        const double sj = ( cw != 0.0 && dA[o] > 0.0 ) ? cw / dA[o] : 0.0;
        const double c = alpha * sd[o];
        double * y = Y + ldY * o;
        for( int k = 0; k < 3; ++k )
        {
            double t = xp[3 * j + k];
            if( cw != 0.0 ) { t += sj * ( s1 * X[ldX * o + k] ); }
            const double val = c * t;
            y[k] = ( beta == 0.0 ) ? val : val + beta * y[k];
        }
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

    void launch_chol_form_m( const CholTask * tasks, int64_t count, const double * Lin, double * Lout, hipStream_t s )
    {
        if( count <= 0 ) { return; }
        hipLaunchKernelGGL( k_chol_form_m, dim3( (unsigned)count ), dim3( 256 ), 0, s, tasks, Lin, Lout );
    }

    void launch_chol_gather( int64_t n, const int32_t * perm, double alpha, const double * B, int64_t ldB, double * bp, hipStream_t s )
    {
        if( n <= 0 ) { return; }
        hipLaunchKernelGGL( k_chol_gather, dim3( (unsigned)( ( n + 255 ) / 256 ) ), dim3( 256 ), 0, s, n, perm, alpha, B, ldB, bp );
    }

    void launch_chol_fwd_small( const CholDev & F, const CholTask * nodes, int64_t count, int G, int rmax, hipStream_t s )
    {
        if( count <= 0 ) { return; }
        const unsigned nb = (unsigned)( ( count * G + 255 ) / 256 );
        const size_t shm = sizeof(double) * 3 * (size_t)rmax * (size_t)( 256 / G );
        if( G <= 8 )       { hipLaunchKernelGGL( k_chol_fwd_small<8>,  dim3( nb ), dim3( 256 ), shm, s, F, nodes, count, rmax ); }
        else if( G <= 16 ) { hipLaunchKernelGGL( k_chol_fwd_small<16>, dim3( nb ), dim3( 256 ), shm, s, F, nodes, count, rmax ); }
        else if( G <= 32 ) { hipLaunchKernelGGL( k_chol_fwd_small<32>, dim3( nb ), dim3( 256 ), shm, s, F, nodes, count, rmax ); }
        else               { hipLaunchKernelGGL( k_chol_fwd_small<64>, dim3( nb ), dim3( 256 ), shm, s, F, nodes, count, rmax ); }
    }

    void launch_chol_fwd_big( const CholDev & F, const CholTask * tasks, int64_t count, int wmax, hipStream_t s )
    {
        if( count <= 0 ) { return; }
        // waves per task by the level's widest supernode: the columns split over 16 waves on the wide
        // levels (more than CHOL_NARROW_WIDTH columns), over 4, 2 or 1 on the others (at most 32 columns
        // per wave), so that narrow levels launch few wavefronts
        if( wmax > CHOL_NARROW_WIDTH ) { hipLaunchKernelGGL( (k_chol_fwd_big<16>), dim3( (unsigned)count ), dim3( 1024 ), 0, s, F, tasks ); }
        else if( wmax > 64 )           { hipLaunchKernelGGL( (k_chol_fwd_big<4>),  dim3( (unsigned)count ), dim3( 256 ),  0, s, F, tasks ); }
        else if( wmax > 32 )           { hipLaunchKernelGGL( (k_chol_fwd_big<2>),  dim3( (unsigned)count ), dim3( 128 ),  0, s, F, tasks ); }
        else                           { hipLaunchKernelGGL( (k_chol_fwd_big<1>),  dim3( (unsigned)count ), dim3( 64 ),   0, s, F, tasks ); }
    }

    void launch_chol_fwd_split( const CholDev & F, const CholTask * front_tasks, int64_t front_count,
                                const CholTask * tasks, int64_t count, double * fr, hipStream_t s )
    {
        if( front_count <= 0 || count <= 0 ) { return; }
        hipLaunchKernelGGL( k_chol_fwd_front, dim3( (unsigned)front_count ), dim3( 64 ), 0, s, F, front_tasks, fr );
        hipLaunchKernelGGL( k_chol_fwd_split, dim3( (unsigned)count ), dim3( 1024 ), 0, s, F, tasks, fr );
    }

    void launch_chol_bwd_small( const CholDev & F, const CholTask * nodes, int64_t count, int G, int rmax, hipStream_t s )
    {
        if( count <= 0 ) { return; }
        const unsigned nb = (unsigned)( ( count * G + 255 ) / 256 );
        const size_t shm = sizeof(double) * 3 * (size_t)rmax * (size_t)( 256 / G );
        if( G <= 8 )       { hipLaunchKernelGGL( k_chol_bwd_small<8>,  dim3( nb ), dim3( 256 ), shm, s, F, nodes, count, rmax ); }
        else if( G <= 16 ) { hipLaunchKernelGGL( k_chol_bwd_small<16>, dim3( nb ), dim3( 256 ), shm, s, F, nodes, count, rmax ); }
        else if( G <= 32 ) { hipLaunchKernelGGL( k_chol_bwd_small<32>, dim3( nb ), dim3( 256 ), shm, s, F, nodes, count, rmax ); }
        else               { hipLaunchKernelGGL( k_chol_bwd_small<64>, dim3( nb ), dim3( 256 ), shm, s, F, nodes, count, rmax ); }
    }

    void launch_chol_bwd_big( const CholDev & F, const CholTask * tasks, int64_t count, int rmax, bool wide, hipStream_t s )
    {
        if( count <= 0 ) { return; }
        if( wide )
        {
            // one rule for launch and kernel: a task stages v at once iff r - k <= lds_rows; levels with
            // longer supernodes get CHOL_BWD_ONE rows, enough for the chunks of CHOL_QC rows of the others
            const int lds_rows = rmax <= CHOL_BWD_ONE ? std::max( rmax, 1 ) : CHOL_BWD_ONE;
            static_assert( CHOL_BWD_ONE >= CHOL_QC, "the staging buffer must hold a chunk" );
            const size_t shm = sizeof(double) * 3 * (size_t)lds_rows;
            hipLaunchKernelGGL( k_chol_bwd_big<true>, dim3( (unsigned)count ), dim3( 256 ), shm, s, F, tasks, lds_rows );
        }
        else { hipLaunchKernelGGL( k_chol_bwd_big<false>, dim3( (unsigned)count ), dim3( 256 ), sizeof(double) * 3 * CHOL_QC, s, F, tasks, CHOL_QC ); }
    }

    void launch_chol_transpose( const CholTask * desc, int64_t count, const double * vals, double * vals_t, hipStream_t s )
    {
        if( count <= 0 ) { return; }
        hipLaunchKernelGGL( k_chol_transpose, dim3( (unsigned)count ), dim3( 256 ), 0, s, desc, count, vals, vals_t );
    }

    void launch_chol_gather_scaled( int64_t n, const int32_t * perm, const double * sd, const double * B, int64_t ldB, double * bp, hipStream_t s )
    {
        if( n <= 0 ) { return; }
        hipLaunchKernelGGL( k_chol_gather_scaled, dim3( (unsigned)( ( n + 255 ) / 256 ) ), dim3( 256 ), 0, s, n, perm, sd, B, ldB, bp );
    }

    void launch_chol_scatter_scaled( int64_t n, const int32_t * perm, const double * xp, const double * sd, const double * X, int64_t ldX,
                                     double cw, const double * dA, double alpha, double beta, double * Y, int64_t ldY, hipStream_t s )
    {
        if( n <= 0 ) { return; }
        hipLaunchKernelGGL( k_chol_scatter_scaled, dim3( (unsigned)( ( n + 255 ) / 256 ) ), dim3( 256 ), 0, s, n, perm, xp, sd, X, ldX, cw, dA, alpha, beta, Y, ldY );
    }

    void launch_chol_scatter( int64_t n, const int32_t * perm, const double * xp, double beta, double * Y, int64_t ldY, hipStream_t s )
    {
        if( n <= 0 ) { return; }
        hipLaunchKernelGGL( k_chol_scatter, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, perm, xp, beta, Y, ldY );
    }

} // namespace rp
