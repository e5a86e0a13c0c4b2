// Very-near field (VF): adaptive subdivision quadrature for the primitive pairs that are neither
// adjacent nor separated by the near-field MAC.
//
//   k_vf_pairs          FMM_Kernel_VF_Common::Compute (src/FMM/FMM_Kernel_VF_Common.hpp:3-116), one lane per
//                       level-1 sub-pair: the reference's traversal starts at level 1 and visits the
//                       (child of S, child of T) pairs in order, so each level-1 sub-pair roots an
//                       independent depth-first subtree.
//                       SimplexHierarchy ToChild/ToParent (src/SimplexHierarchy/SimplexHierarchy_1D.hpp, _2D.hpp),
//                       TP0_Kernel_VF::compute (src/Energies/TangentPointEnergy0/TP0_Kernel_VF.hpp:158-338),
//                       PseudoLaplacian_Kernel_VF::compute (src/Metrics/PseudoLaplacian/PseudoLaplacian_Kernel_VF.hpp:135-176).
//                       The GJK MAC (src/GJK.hpp:1162-1172) is evaluated with the exact simplex distance.
//   k_vf_rows           FMM_Kernel_VF::writeS (FMM_Kernel_VF.hpp:285-305), VF_Accumulator AddReduce and
//                       FillLowerTriangleFromUpperTriangle (TP0_Traversor.hpp:296-307)
//   k_vf_spmm_post      TP0_Traversor::VF_MultiplyMetric (TP0_Traversor.hpp:453-478) followed by the post operator
//   k_vf_pl_spmm_post   PseudoLaplacian_Traversor::VF_MultiplyMetric (PseudoLaplacian_Traversor.hpp:334-360)
//
// Symmetric pattern, no atomics: every entry (i, j) of the full pattern is traversed with S = i and
// produces the row-side (S) terms; the T-side terms of the reference are the S-side terms of the
// transposed entry. The energy counts upper entries only; the metric values of a lower entry are
// copied from its transposed upper entry.

#include "rp_launch.hpp"
#include "k_pair.hpp"

namespace rp
{
    namespace
    {
        inline dim3 vf_grid( int64_t n ) { return dim3( (unsigned)( ( n + 255 ) / 256 ) ); }

        struct D3 { double x, y, z; };
        __device__ __forceinline__ D3 sub3( D3 a, D3 b ) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
        __device__ __forceinline__ D3 add3( D3 a, D3 b ) { return { a.x + b.x, a.y + b.y, a.z + b.z }; }
        __device__ __forceinline__ D3 mul3( double s, D3 a ) { return { s * a.x, s * a.y, s * a.z }; }
        __device__ __forceinline__ double dot3( D3 a, D3 b ) { return a.x * b.x + a.y * b.y + a.z * b.z; }
        __device__ __forceinline__ double clamp01( double t ) { return t < 0 ? 0 : (t > 1 ? 1 : t); }
        __device__ __forceinline__ double dmin( double a, double b ) { return (b < a) ? b : a; }

        // Closest point on triangle abc to p (Voronoi regions of vertices, edges, face).
        __device__ D3 closest_on_triangle( D3 p, D3 a, D3 b, D3 c )
        {
            _Pragma("clang fp contract(off)")
            const D3 ab = sub3(b, a), ac = sub3(c, a), ap = sub3(p, a);
            const double d1 = dot3(ab, ap), d2 = dot3(ac, ap);
            if( d1 <= 0 && d2 <= 0 ) { return a; }
            const D3 bp = sub3(p, b);
            const double d3 = dot3(ab, bp), d4 = dot3(ac, bp);
            if( d3 >= 0 && d4 <= d3 ) { return b; }
            const double vc = d1 * d4 - d3 * d2;
            if( vc <= 0 && d1 >= 0 && d3 <= 0 ) { return add3( a, mul3( d1 / (d1 - d3), ab ) ); }
            const D3 cp = sub3(p, c);
            const double d5 = dot3(ab, cp), d6 = dot3(ac, cp);
            if( d6 >= 0 && d5 <= d6 ) { return c; }
            const double vb = d5 * d2 - d1 * d6;
            if( vb <= 0 && d2 >= 0 && d6 <= 0 ) { return add3( a, mul3( d2 / (d2 - d6), ac ) ); }
            const double va = d3 * d6 - d5 * d4;
            if( va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0 ) { return add3( b, mul3( (d4 - d3) / ((d4 - d3) + (d5 - d6)), sub3(c, b) ) ); }
            const double den = 1.0 / (va + vb + vc);
            return add3( a, add3( mul3( vb * den, ab ), mul3( vc * den, ac ) ) );
        }

        // Squared distance of the segments [p1,q1] and [p2,q2].
        __device__ double segments_d2( D3 p1, D3 q1, D3 p2, D3 q2 )
        {
            _Pragma("clang fp contract(off)")
            const D3 d1 = sub3(q1, p1), d2 = sub3(q2, p2), r = sub3(p1, p2);
            const double a = dot3(d1, d1), e = dot3(d2, d2), f = dot3(d2, r);
            double s = 0, t = 0;
            if( a <= 1e-300 && e <= 1e-300 ) { return dot3(r, r); }
            if( a <= 1e-300 ) { t = clamp01( f / e ); }
            else
            {
                const double c = dot3(d1, r);
                if( e <= 1e-300 ) { s = clamp01( -c / a ); }
                else
                {
                    const double b = dot3(d1, d2), den = a * e - b * b;
                    s = (den != 0) ? clamp01( (b * f - c * e) / den ) : 0.0;
                    t = (b * s + f) / e;
                    if( t < 0 ) { t = 0; s = clamp01( -c / a ); }
                    else if( t > 1 ) { t = 1; s = clamp01( (b - c) / a ); }
                }
            }
            const D3 w = sub3( add3( p1, mul3( s, d1 ) ), add3( p2, mul3( t, d2 ) ) );
            return dot3(w, w);
        }

        // Squared distance of two disjoint simplices (N = 1: segments, N = 2: triangles).
        template<int N>
        __device__ double simplex_d2( const D3 * P, const D3 * Q )
        {
            if constexpr ( N == 1 ) { return segments_d2( P[0], P[1], Q[0], Q[1] ); }
            else
            {
                _Pragma("clang fp contract(off)")
                double d2 = 1e300;
                #pragma unroll
                for( int i = 0; i < 3; ++i )
                {
                    D3 c = closest_on_triangle( P[i], Q[0], Q[1], Q[2] );
                    d2 = dmin( d2, dot3( sub3(P[i], c), sub3(P[i], c) ) );
                    c = closest_on_triangle( Q[i], P[0], P[1], P[2] );
                    d2 = dmin( d2, dot3( sub3(Q[i], c), sub3(Q[i], c) ) );
                }
                #pragma unroll
                for( int i = 0; i < 3; ++i )
                {
                    #pragma unroll
                    for( int j = 0; j < 3; ++j ) { d2 = dmin( d2, segments_d2( P[i], P[(i+1)%3], Q[j], Q[(j+1)%3] ) ); }
                }
                return d2;
            }
        }

        // Sub-simplex of one primitive: barycentric corners and center, level, child ids, path code
        // (the state of the reference's SimplexHierarchy, kept in registers; the child index is
        // resolved by selects so that nothing is indexed dynamically).
        template<int N>
        struct Sub
        {
            static constexpr int S  = N + 1;
            static constexpr int NC = (N == 2) ? 4 : 2;
            double c[S][S];
            double m[S];
            int level, child, former;
            uint64_t column;
            double scale, weight;

            __device__ __forceinline__ void reset()
            {
                #pragma unroll
                for( int i = 0; i < S; ++i )
                {
                    m[i] = 1.0 / S;
                    #pragma unroll
                    for( int j = 0; j < S; ++j ) { c[i][j] = (i == j) ? 1.0 : 0.0; }
                }
                level = 0; child = 0; former = -1; column = 0; scale = 1.0; weight = 1.0;
            }

            __device__ __forceinline__ void move( int k, double f )   // f = 0.5: to child k, f = 2: back from child k
            {
                _Pragma("clang fp contract(off)")
                if( k < S )
                {
                    double ck[S];
                    #pragma unroll
                    for( int j = 0; j < S; ++j )
                    {
                        ck[j] = c[0][j];
                        #pragma unroll
                        for( int i = 1; i < S; ++i ) { ck[j] = (k == i) ? c[i][j] : ck[j]; }
                    }
                    #pragma unroll
                    for( int i = 0; i < S; ++i )
                    {
                        #pragma unroll
                        for( int j = 0; j < S; ++j )
                        {
                            if( i != k ) { c[i][j] = ck[j] + f * (c[i][j] - ck[j]); }
                            else         { m[j] = ck[j] + f * (m[j] - ck[j]); }
                        }
                    }
                }
                else
                {
                    #pragma unroll
                    for( int i = 0; i < S; ++i )
                    {
                        #pragma unroll
                        for( int j = 0; j < S; ++j ) { c[i][j] = m[j] - f * (c[i][j] - m[j]); }
                    }
                }
            }

            __device__ __forceinline__ void to_child( int k )
            {
                former = child; child = k; column = (uint64_t)NC * column + (uint64_t)k; ++level;
                scale *= 0.5; weight *= (N == 2 ? 0.25 : 0.5);
                move( k, 0.5 );
            }

            __device__ __forceinline__ void to_parent()
            {
                const int k = child;
                column = (column - (uint64_t)child) / (uint64_t)NC;
                former = child; child = (int)(column % (uint64_t)NC); --level;
                scale *= 2.0; weight *= (N == 2 ? 4.0 : 2.0);
                move( k, 2.0 );
            }

            __device__ __forceinline__ void vertices( const D3 * hull, D3 * out ) const
            {
                _Pragma("clang fp contract(off)")
                #pragma unroll
                for( int i = 0; i < S; ++i )
                {
                    D3 p = mul3( c[i][0], hull[0] );
                    #pragma unroll
                    for( int j = 1; j < S; ++j ) { p = add3( p, mul3( c[i][j], hull[j] ) ); }
                    out[i] = p;
                }
            }

            __device__ __forceinline__ D3 point( const D3 * hull ) const
            {
                D3 p = mul3( m[0], hull[0] );
                #pragma unroll
                for( int j = 1; j < S; ++j ) { p = add3( p, mul3( m[j], hull[j] ) ); }
                return p;
            }
        };

        // Max squared vertex distance from the vertex average (Polytope::FromCoordinates,
        // src/Primitives/Polytope.hpp:69-111).
        template<int S>
        __device__ double prim_r2( const D3 * h )
        {
            _Pragma("clang fp contract(off)")
            double cx = h[0].x, cy = h[0].y, cz = h[0].z;
            #pragma unroll
            for( int j = 1; j < S; ++j ) { cx += h[j].x; cy += h[j].y; cz += h[j].z; }
            cx *= 1.0 / S; cy *= 1.0 / S; cz *= 1.0 / S;
            double r2 = 0;
            #pragma unroll
            for( int j = 0; j < S; ++j )
            {
                double d = h[j].x - cx;
                double sq = d * d;
                d = h[j].y - cy; sq += d * d;
                d = h[j].z - cz; sq += d * d;
                r2 = (sq > r2) ? sq : r2;
            }
            return r2;
        }

        template<int S>
        __device__ __forceinline__ void load_hull( const VFArgs & A, int32_t prim, D3 * h )
        {
            const int64_t t = A.P_ordering[prim];
            #pragma unroll
            for( int l = 0; l < S; ++l )
            {
                const int64_t v = A.simplices[S * t + l];
                h[l] = { A.coords[3 * v], A.coords[3 * v + 1], A.coords[3 * v + 2] };
            }
        }
    }

    // Record per pattern entry (VF_REC doubles). Tangent-point: [0, ND) S-side moment derivatives,
    // [16] energy (upper entries), [17..18] (a0, a1), [19..20] row-diagonal terms. Pseudo-Laplacian:
    // [0] value, [1] row-diagonal term. Both: [21] evaluations, [22] deepest level, [23] 1 if the
    // evaluation limit stopped the traversal.
    template<int N, int MODE>
    __global__ __launch_bounds__(256) void k_vf_pairs( VFArgs A )
    {
        constexpr int S = N + 1, NC = (N == 2) ? 4 : 2, G = NC * NC;
        constexpr int ND = 1 + 3 * S + 6;
        const int64_t gid = (int64_t)blockIdx.x * 256 + threadIdx.x;
        const int64_t e = gid / G;
        const int sub = (int)(gid % G);
        double acc[VF_REC];
        #pragma unroll
        for( int k = 0; k < VF_REC; ++k ) { acc[k] = 0.0; }
        if( e < A.nnz )
        {
            const int32_t i = A.row[e], j = A.col[e];
            D3 hs[S], ht[S];
            load_hull<S>( A, i, hs );
            load_hull<S>( A, j, ht );
            const double r2S = prim_r2<S>( hs ), r2T = prim_r2<S>( ht );
            const double a = A.a[i], b = A.a[j];
            double P[6], Q[6];
            #pragma unroll
            for( int k = 0; k < 6; ++k ) { P[k] = A.P[k * A.T + i]; Q[k] = A.P[k * A.T + j]; }
            const bool upper = j > i;
            const double b_over_a = b / a;
            Sub<N> X, Y;
            X.reset(); Y.reset();
            X.to_child( sub / NC ); Y.to_child( sub % NC );
            bool from_above = true;
            int64_t evals = 0;
            int deepest = 0;
            while( true )
            {
                if( from_above )
                {
                    bool leaf = X.level >= A.max_ref;
                    if( !leaf )
                    {
                        D3 PS[S], PT[S];
                        X.vertices( hs, PS ); Y.vertices( ht, PT );
                        const double rx = X.scale * X.scale * r2S, ry = Y.scale * Y.scale * r2T;
                        const double rr = (rx < ry) ? ry : rx;
                        leaf = rr < A.theta2 * simplex_d2<N>( PS, PT );
                    }
                    if( leaf )
                    {
                        if( evals >= A.cap ) { acc[23] = 1.0; break; }
                        ++evals;
                        deepest = max( deepest, min( X.level, A.max_ref ) );
                        const double w = X.weight * Y.weight;
                        const D3 x = X.point( hs ), y = Y.point( ht );
                        const double v[3] = { y.x - x.x, y.y - x.y, y.z - x.z };
                        if constexpr ( MODE == 0 )
                        {
                            PairTerms<double> t;
                            pair_terms<double, KIND_GENERIC, true, true, true>( A.ex, v, P, Q, t );
                            if( upper && A.energy ) { acc[16] += a * t.E * b; }    // no quadrature weight, as TP0_Kernel_VF.hpp:273
                            if( A.diff )
                            {
                                const double wb = w * b;
                                const double dx = t.g[0] * x.x + t.g[1] * x.y + t.g[2] * x.z;
                                acc[0] += wb * ( t.E0 + dx );
                                #pragma unroll
                                for( int h = 0; h < S; ++h )
                                {
                                    acc[1 + 3 * h + 0] -= X.m[h] * wb * t.g[0];
                                    acc[1 + 3 * h + 1] -= X.m[h] * wb * t.g[1];
                                    acc[1 + 3 * h + 2] -= X.m[h] * wb * t.g[2];
                                }
                                #pragma unroll
                                for( int k = 0; k < 6; ++k ) { acc[1 + 3 * S + k] += wb * t.hK * t.Vk[k]; }
                            }
                            if( A.metric )
                            {
                                const double a1 = w * t.a1;
                                const double a0 = w * ( w * t.a0 );        // the weight enters twice, as TP0_Kernel_VF.hpp:320
                                acc[17] += a0; acc[18] += a1;
                                acc[19] -= b_over_a * a0; acc[20] -= b_over_a * a1;
                            }
                        }
                        else
                        {
                            const double r2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
                            const double val = -w * pow( r2, A.pl_sexp );
                            acc[0] += val;
                            acc[1] -= b_over_a * val;
                        }
                        X.to_parent(); Y.to_parent();
                        from_above = false;
                    }
                    else { X.to_child( 0 ); Y.to_child( 0 ); }
                }
                else
                {
                    if( X.level == 0 ) { break; }
                    const int sk = X.former, tk = Y.former;
                    if( tk + 1 < NC )      { X.to_child( sk );     Y.to_child( tk + 1 ); from_above = true; }
                    else if( sk + 1 < NC ) { X.to_child( sk + 1 ); Y.to_child( 0 );      from_above = true; }
                    else                   { X.to_parent();        Y.to_parent(); }
                }
            }
            acc[21] = upper ? (double)evals : 0.0;      // the reference traverses upper entries only
            acc[22] = (double)deepest;
        }
        (void)ND;
        // sum over the G lanes of the entry (groups of G aligned lanes inside one wavefront)
        #pragma unroll
        for( int k = 0; k < VF_REC; ++k )
        {
            double vk = acc[k];
            #pragma unroll
            for( int off = G / 2; off > 0; off >>= 1 )
            {
                const double o = __shfl_xor( vk, off );
                vk = ( k >= 22 ) ? fmax( vk, o ) : vk + o;
            }
            acc[k] = vk;
        }
        if( e < A.nnz && sub == 0 )
        {
            double * r = A.rec + (int64_t)VF_REC * e;
            #pragma unroll
            for( int k = 0; k < VF_REC; ++k ) { r[k] = acc[k]; }
        }
    }

    // Per VF row (in CSR order of its entries): moment derivatives into dx (10 x T, [0] and the
    // projector rows 4..9) and vf_dh (T x 3(n+1), hull-vertex rows, written), row-diagonal terms,
    // metric values of the full pattern, and the row's energy partial.
    template<int N, int MODE>
    __global__ __launch_bounds__(256) void k_vf_rows( VFRowArgs A )
    {
        constexpr int S = N + 1;
        const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
        if( r >= A.nrows ) { return; }
        const int32_t i = A.rows[r];
        const int64_t kb = A.rp[i], ke = A.rp[i + 1], kd = A.dg[i];
        if constexpr ( MODE == 0 )
        {
            double DX[1 + 3 * S + 6];
            #pragma unroll
            for( int h = 0; h < 1 + 3 * S + 6; ++h ) { DX[h] = 0.0; }
            double en = 0.0, d0 = 0.0, d1 = 0.0;
            for( int64_t k = kb; k < ke; ++k )
            {
                const double * rc = A.rec + (int64_t)VF_REC * k;
                #pragma unroll
                for( int h = 0; h < 1 + 3 * S + 6; ++h ) { DX[h] += rc[h]; }
                if( k >= kd ) { en += rc[16]; }
                d0 += rc[19]; d1 += rc[20];
                if( A.vals )
                {
                    const double * ru = A.rec + (int64_t)VF_REC * A.tr[k];
                    A.vals[2 * k] = ru[17]; A.vals[2 * k + 1] = ru[18];
                }
            }
            if( A.dx )
            {
                A.dx[i] += DX[0];
                #pragma unroll
                for( int k = 0; k < 6; ++k ) { A.dx[(4 + k) * A.T + i] += DX[1 + 3 * S + k]; }
                #pragma unroll
                for( int h = 0; h < 3 * S; ++h ) { A.dh[(int64_t)3 * S * i + h] = DX[1 + h]; }
            }
            if( A.diag ) { A.diag[2 * (int64_t)i] = d0; A.diag[2 * (int64_t)i + 1] = d1; }
            if( A.epart ) { A.epart[r] = en; }
        }
        else
        {
            double d = 0.0;
            for( int64_t k = kb; k < ke; ++k )
            {
                d += A.rec[(int64_t)VF_REC * k + 1];
                A.vals[k] = A.rec[(int64_t)VF_REC * A.tr[k]];
            }
            A.diag[i] = d;
        }
    }

    // Evaluation count (sum), deepest level and limit flag (max) over all entries.
    __global__ __launch_bounds__(256) void k_vf_stats( int64_t nnz, const double * rec, double * out )
    {
        __shared__ double sh[3][256];
        double s = 0.0, mx = 0.0, cap = 0.0;
        for( int64_t e = threadIdx.x; e < nnz; e += 256 )
        {
            const double * r = rec + (int64_t)VF_REC * e;
            s += r[21]; mx = fmax( mx, r[22] ); cap = fmax( cap, r[23] );
        }
        sh[0][threadIdx.x] = s; sh[1][threadIdx.x] = mx; sh[2][threadIdx.x] = cap;
        __syncthreads();
        for( int w = 128; w > 0; w >>= 1 )
        {
            if( (int)threadIdx.x < w )
            {
                sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
                sh[1][threadIdx.x] = fmax( sh[1][threadIdx.x], sh[1][threadIdx.x + w] );
                sh[2][threadIdx.x] = fmax( sh[2][threadIdx.x], sh[2][threadIdx.x + w] );
            }
            __syncthreads();
        }
        if( threadIdx.x == 0 ) { out[0] = sh[0][0]; out[1] = sh[1][0]; out[2] = sh[2][0]; }
    }

    // Metric product over the VF rows: y_i = diag_i x_i + sum_j (a0, a1)_ij x_j on the 4 x 3 mixed
    // primitive vectors, then the post operator added into the vertex-slot output S_out.
    template<int SIZE>
    __global__ __launch_bounds__(256) void k_vf_spmm_post( int64_t nrows, const int32_t * rows, const int64_t * rp, const int32_t * ci,
                                                           const double * vals, const double * diag, const double * P_in,
                                                           const double * pre_vals, double * S_out )
    {
        const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
        if( r >= nrows ) { return; }
        const int32_t i = rows[r];
        double y[12];
        const double * xi = P_in + 12 * (int64_t)i;
        #pragma unroll
        for( int h = 0; h < 3; ++h ) { y[h] = diag[2 * (int64_t)i] * xi[h]; }
        #pragma unroll
        for( int h = 3; h < 12; ++h ) { y[h] = diag[2 * (int64_t)i + 1] * xi[h]; }
        for( int64_t k = rp[i]; k < rp[i + 1]; ++k )
        {
            const double * x = P_in + 12 * (int64_t)ci[k];
            const double a0 = vals[2 * k], a1 = vals[2 * k + 1];
            #pragma unroll
            for( int h = 0; h < 3; ++h ) { y[h] += a0 * x[h]; }
            #pragma unroll
            for( int h = 3; h < 12; ++h ) { y[h] += a1 * x[h]; }
        }
        const double * pv = pre_vals + 4 * SIZE * (int64_t)i;
        #pragma unroll
        for( int sl = 0; sl < SIZE; ++sl )
        {
            #pragma unroll
            for( int cc = 0; cc < 3; ++cc )
            {
                double t = 0.0;
                #pragma unroll
                for( int rr = 0; rr < 4; ++rr ) { t += pv[SIZE * rr + sl] * y[3 * rr + cc]; }
                S_out[(SIZE * (int64_t)i + sl) * 3 + cc] += t;
            }
        }
    }

    // Pseudo-Laplacian (LowOrder) product over the VF rows, added into S_out.
    template<int SIZE>
    __global__ __launch_bounds__(256) void k_vf_pl_spmm_post( int64_t nrows, const int32_t * rows, const int64_t * rp, const int32_t * ci,
                                                              const double * vals, const double * diag, const double * P_in,
                                                              const double * pre_vals, double * S_out )
    {
        const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
        if( r >= nrows ) { return; }
        const int32_t i = rows[r];
        double y[3];
        #pragma unroll
        for( int h = 0; h < 3; ++h ) { y[h] = diag[i] * P_in[3 * (int64_t)i + h]; }
        for( int64_t k = rp[i]; k < rp[i + 1]; ++k )
        {
            #pragma unroll
            for( int h = 0; h < 3; ++h ) { y[h] += vals[k] * P_in[3 * (int64_t)ci[k] + h]; }
        }
        const double * pv = pre_vals + 4 * SIZE * (int64_t)i;
        #pragma unroll
        for( int sl = 0; sl < SIZE; ++sl )
        {
            #pragma unroll
            for( int h = 0; h < 3; ++h ) { S_out[(SIZE * (int64_t)i + sl) * 3 + h] += pv[sl] * y[h]; }
        }
    }

    // ------------------------------------------------------------------ launchers
    void launch_vf_pairs( int n, int mode, const VFArgs & A, hipStream_t s )
    {
        if( A.nnz <= 0 ) { return; }
        const int64_t G = (n == 2) ? 16 : 4;
        const dim3 grid = vf_grid( A.nnz * G );
        if( n == 2 ) { if( mode == 0 ) { hipLaunchKernelGGL( (k_vf_pairs<2,0>), grid, dim3(256), 0, s, A ); } else { hipLaunchKernelGGL( (k_vf_pairs<2,1>), grid, dim3(256), 0, s, A ); } }
        else         { if( mode == 0 ) { hipLaunchKernelGGL( (k_vf_pairs<1,0>), grid, dim3(256), 0, s, A ); } else { hipLaunchKernelGGL( (k_vf_pairs<1,1>), grid, dim3(256), 0, s, A ); } }
    }

    void launch_vf_rows( int n, int mode, const VFRowArgs & A, hipStream_t s )
    {
        if( A.nrows <= 0 ) { return; }
        const dim3 grid = vf_grid( A.nrows );
        if( n == 2 ) { if( mode == 0 ) { hipLaunchKernelGGL( (k_vf_rows<2,0>), grid, dim3(256), 0, s, A ); } else { hipLaunchKernelGGL( (k_vf_rows<2,1>), grid, dim3(256), 0, s, A ); } }
        else         { if( mode == 0 ) { hipLaunchKernelGGL( (k_vf_rows<1,0>), grid, dim3(256), 0, s, A ); } else { hipLaunchKernelGGL( (k_vf_rows<1,1>), grid, dim3(256), 0, s, A ); } }
    }

    void launch_vf_stats( int64_t nnz, const double * rec, double * out, hipStream_t s )
    {
        hipLaunchKernelGGL( k_vf_stats, dim3(1), dim3(256), 0, s, nnz, rec, out );
    }

    void launch_vf_spmm_post( int size, int64_t nrows, const int32_t * rows, const int64_t * rp, const int32_t * ci, const double * vals,
                              const double * diag, const double * P_in, const double * pre_vals, double * S_out, hipStream_t s )
    {
        if( nrows <= 0 ) { return; }
        if( size == 3 ) { hipLaunchKernelGGL( (k_vf_spmm_post<3>), vf_grid( nrows ), dim3(256), 0, s, nrows, rows, rp, ci, vals, diag, P_in, pre_vals, S_out ); }
        else            { hipLaunchKernelGGL( (k_vf_spmm_post<2>), vf_grid( nrows ), dim3(256), 0, s, nrows, rows, rp, ci, vals, diag, P_in, pre_vals, S_out ); }
    }

    void launch_vf_pl_spmm_post( int size, int64_t nrows, const int32_t * rows, const int64_t * rp, const int32_t * ci, const double * vals,
                                 const double * diag, const double * P_in, const double * pre_vals, double * S_out, hipStream_t s )
    {
        if( nrows <= 0 ) { return; }
        if( size == 3 ) { hipLaunchKernelGGL( (k_vf_pl_spmm_post<3>), vf_grid( nrows ), dim3(256), 0, s, nrows, rows, rp, ci, vals, diag, P_in, pre_vals, S_out ); }
        else            { hipLaunchKernelGGL( (k_vf_pl_spmm_post<2>), vf_grid( nrows ), dim3(256), 0, s, nrows, rows, rp, ci, vals, diag, P_in, pre_vals, S_out ); }
    }
}
