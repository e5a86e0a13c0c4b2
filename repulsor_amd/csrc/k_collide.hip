// Maximal safe step size of a moving mesh (the line-search bound after the metric solve):
//   SimplicialMesh::MaximumSafeStepSize     src/SimplicialMesh.hpp:441-468
//   CollisionTree::MaximumSafeStepSize      src/CollisionTree.hpp (traversal, minimum over the kernels)
//   MaximumSafeStepSize_Kernel              src/CollisionTree/MaximumSafeStepSize_Kernel.hpp (adjacent pairs skipped)
//   CollisionFinder::FindMaximumSafeStepSize src/CollisionFinder.hpp (bisection of [0, t_init])
//   GJK::IntersectingQ on MovingPolytopes   src/GJK.hpp:618-797, src/Primitives/MovingPolytope.hpp
//
//   k_prim_swept_boxes   axis-aligned box of each simplex over [0, t_max] (its points at 0 and t_max)
//   k_leaf_boxes, k_level_boxes   the same for clusters (leaves, then one depth level per launch)
//   k_traverse_step      one level of the dual-tree traversal: cluster pairs whose boxes meet are split
//                        (a cluster with itself into its three child pairs, two internal clusters into
//                        four, a leaf with an internal cluster into two); leaf pairs are emitted
//   k_leaf_pair_steps    one workgroup per leaf pair, one lane per simplex pair: skip adjacent pairs and
//                        pairs whose boxes do not meet, else the reference's bisection with a GJK test of
//                        the space-time hulls (points of the simplex at the interval's two ends, lifted by
//                        time); the minimum over all pairs by an atomic minimum on the double's bits
//                        (non-negative doubles order like their bit patterns), so the result does not
//                        depend on the order.
// Each pair starts its bisection at t_init = t_max (the reference starts later pairs of a thread at its
// running minimum; that changes results only within the tolerance).

#include "rp_launch.hpp"

namespace rp
{
    namespace
    {
        inline dim3 col_grid( int64_t n ) { return dim3( (unsigned)( ( n + 255 ) / 256 ) ); }

        __device__ __forceinline__ bool boxes_meet( const double * a, const double * b )
        {
            return a[0] <= b[3] && b[0] <= a[3] && a[1] <= b[4] && b[1] <= a[4] && a[2] <= b[5] && b[2] <= a[5];
        }

        __device__ __forceinline__ double dot3( const double * a, const double * b ) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

        // Moving simplex data, per primitive in cluster order: for each of its M vertices the position
        // (3 doubles) and then the velocity (3 doubles).
        //
        // The space-time hulls of the two simplices over [a, b] (their points at a and at b, lifted by
        // time) meet exactly when the origin lies in the convex hull of the 3D differences
        //     D = { p_i(a) - q_j(a) } u { p_i(b) - q_j(b) }:
        // a common point has equal time coordinates, so both sides put the same weight s on their points
        // at b, and (1 - s) A_P + s B_P meets (1 - s) A_Q + s B_Q for hulls A, B of the points at a and b.
        // GJK therefore runs in 3D on D (at most 2 M^2 points, never formed), with the reference's
        // tolerance (eps^2 times the smaller squared radius of the two space-time hulls) and stopping rules.

        // argmin over D of v . x: returns v . w and writes w
        template<int M>
        __device__ __forceinline__ double support3( const double * P, const double * Q, double a, double b, const double v[3], double w[3] )
        {
            double vp[M], vu[M], vq[M], vv[M];
            #pragma unroll
            for( int i = 0; i < M; ++i )
            {
                vp[i] = dot3( v, P + 6 * i ); vu[i] = dot3( v, P + 6 * i + 3 );
                vq[i] = dot3( v, Q + 6 * i ); vv[i] = dot3( v, Q + 6 * i + 3 );
            }
            double best = 0.0, bt = a;
            int bi = 0, bj = 0;
            #pragma unroll
            for( int s = 0; s < 2; ++s )
            {
                const double t = s == 0 ? a : b;
                double mp = vp[0] + t * vu[0], mq = vq[0] + t * vv[0];
                int ip = 0, jq = 0;
                #pragma unroll
                for( int i = 1; i < M; ++i )
                {
                    const double x = vp[i] + t * vu[i];
                    if( x < mp ) { mp = x; ip = i; }
                    const double y = vq[i] + t * vv[i];
                    if( y > mq ) { mq = y; jq = i; }
                }
                const double val = mp - mq;
                if( s == 0 || val < best ) { best = val; bi = ip; bj = jq; bt = t; }
            }
            const double * pb = P + 6 * bi, * qb = Q + 6 * bj;
            #pragma unroll
            for( int c = 0; c < 3; ++c ) { w[c] = ( pb[c] + bt * pb[3 + c] ) - ( qb[c] + bt * qb[3 + c] ); }
            return best;
        }

        // Affine minimum-norm point of segment {W0, B}, if it lies strictly inside.
        __device__ __forceinline__ void try_segment( const double W0[3], const double B[3], int id, double & best, int & best_id, double v[3] )
        {
            const double e[3] = { B[0] - W0[0], B[1] - W0[1], B[2] - W0[2] };
            const double ee = dot3( e, e );
            if( !( ee > 0.0 ) ) { return; }
            const double t = -dot3( W0, e ) / ee;
            if( !( t > 0.0 && t < 1.0 ) ) { return; }
            double x[3];
            #pragma unroll
            for( int c = 0; c < 3; ++c ) { x[c] = W0[c] + t * e[c]; }
            const double n2 = dot3( x, x );
            if( n2 < best ) { best = n2; best_id = id; v[0] = x[0]; v[1] = x[1]; v[2] = x[2]; }
        }

        // Affine minimum-norm point of triangle {W0, B, C}, if it lies strictly inside.
        __device__ __forceinline__ void try_triangle( const double W0[3], const double B[3], const double C[3], int id, double & best, int & best_id,
                                                      double v[3] )
        {
            const double e1[3] = { B[0] - W0[0], B[1] - W0[1], B[2] - W0[2] };
            const double e2[3] = { C[0] - W0[0], C[1] - W0[1], C[2] - W0[2] };
            const double g11 = dot3( e1, e1 ), g12 = dot3( e1, e2 ), g22 = dot3( e2, e2 );
            const double r1 = -dot3( W0, e1 ), r2 = -dot3( W0, e2 );
            const double det = g11 * g22 - g12 * g12;
            if( !( det > 1e-13 * g11 * g22 ) ) { return; }
            const double m1 = ( r1 * g22 - r2 * g12 ) / det, m2 = ( g11 * r2 - g12 * r1 ) / det;
            if( !( m1 > 0.0 && m2 > 0.0 && 1.0 - m1 - m2 > 0.0 ) ) { return; }
            double x[3];
            #pragma unroll
            for( int c = 0; c < 3; ++c ) { x[c] = W0[c] + m1 * e1[c] + m2 * e2[c]; }
            const double n2 = dot3( x, x );
            if( n2 < best ) { best = n2; best_id = id; v[0] = x[0]; v[1] = x[1]; v[2] = x[2]; }
        }

        // Closest point of conv(W[0..n)) to the origin, over the faces that contain the newest point W[0]
        // (it belongs to the closest face whenever it improved the support value). Subsets are numbered
        // by the bit mask of the other slots they use. W is compacted to the chosen face; returns true
        // when the origin lies inside the tetrahedron.
        __device__ __forceinline__ bool closest_face( double W[4][3], int & n, double v[3] )
        {
            double best = dot3( W[0], W[0] );
            int id = 0;
            v[0] = W[0][0]; v[1] = W[0][1]; v[2] = W[0][2];
            if( n >= 2 ) { try_segment( W[0], W[1], 1, best, id, v ); }
            if( n >= 3 ) { try_segment( W[0], W[2], 2, best, id, v ); try_triangle( W[0], W[1], W[2], 3, best, id, v ); }
            if( n >= 4 )
            {
                try_segment( W[0], W[3], 4, best, id, v );
                try_triangle( W[0], W[1], W[3], 5, best, id, v );
                try_triangle( W[0], W[2], W[3], 6, best, id, v );
                double e[3][3];
                #pragma unroll
                for( int k = 0; k < 3; ++k ) { for( int c = 0; c < 3; ++c ) { e[k][c] = W[k + 1][c] - W[0][c]; } }
                const double c23[3] = { e[1][1] * e[2][2] - e[1][2] * e[2][1], e[1][2] * e[2][0] - e[1][0] * e[2][2], e[1][0] * e[2][1] - e[1][1] * e[2][0] };
                const double c31[3] = { e[2][1] * e[0][2] - e[2][2] * e[0][1], e[2][2] * e[0][0] - e[2][0] * e[0][2], e[2][0] * e[0][1] - e[2][1] * e[0][0] };
                const double c12[3] = { e[0][1] * e[1][2] - e[0][2] * e[1][1], e[0][2] * e[1][0] - e[0][0] * e[1][2], e[0][0] * e[1][1] - e[0][1] * e[1][0] };
                const double det = dot3( e[0], c23 );
                const double scale = sqrt( dot3( e[0], e[0] ) * dot3( e[1], e[1] ) * dot3( e[2], e[2] ) );
                if( fabs( det ) > 1e-13 * scale )
                {
                    const double m1 = -dot3( W[0], c23 ) / det, m2 = -dot3( W[0], c31 ) / det, m3 = -dot3( W[0], c12 ) / det;
                    if( m1 > 0.0 && m2 > 0.0 && m3 > 0.0 && 1.0 - m1 - m2 - m3 > 0.0 ) { v[0] = v[1] = v[2] = 0.0; return true; }
                }
            }
            switch( id )
            {
                case 0: n = 1; break;
                case 1: n = 2; break;
                case 2: for( int c = 0; c < 3; ++c ) { W[1][c] = W[2][c]; } n = 2; break;
                case 3: n = 3; break;
                case 4: for( int c = 0; c < 3; ++c ) { W[1][c] = W[3][c]; } n = 2; break;
                case 5: for( int c = 0; c < 3; ++c ) { W[2][c] = W[3][c]; } n = 3; break;
                default: for( int c = 0; c < 3; ++c ) { W[1][c] = W[2][c]; W[2][c] = W[3][c]; } n = 3; break;
            }
            return false;
        }

        // GJK::IntersectingQ of the two space-time hulls over [a, b] (time lifted by Ts for the radii).
        template<int M>
        __device__ bool hulls_meet( const double * P, const double * Q, double a, double b, double Ts )
        {
            const double tm = 0.5 * ( a + b );
            double cP[3] = {0,0,0}, cQ[3] = {0,0,0};
            #pragma unroll
            for( int i = 0; i < M; ++i )
            {
                #pragma unroll
                for( int c = 0; c < 3; ++c ) { cP[c] += ( P[6 * i + c] + tm * P[6 * i + 3 + c] ) / M; cQ[c] += ( Q[6 * i + c] + tm * Q[6 * i + 3 + c] ) / M; }
            }
            double rP = 0, rQ = 0;
            #pragma unroll
            for( int i = 0; i < M; ++i )
            {
                #pragma unroll
                for( int s = 0; s < 2; ++s )
                {
                    const double t = s == 0 ? a : b;
                    const double z = ( t - tm ) * Ts;
                    double dp = z * z, dq = z * z;
                    #pragma unroll
                    for( int c = 0; c < 3; ++c )
                    {
                        const double x = P[6 * i + c] + t * P[6 * i + 3 + c] - cP[c]; dp += x * x;
                        const double y = Q[6 * i + c] + t * Q[6 * i + 3 + c] - cQ[c]; dq += y * y;
                    }
                    rP = fmax( rP, dp ); rQ = fmax( rQ, dq );
                }
            }
            const double eps2 = 2.220446049250313e-16;   // squared tolerance factor: the double epsilon
            double TOL2 = eps2 * fmin( rP, rQ );
            if( !( TOL2 > 0.0 ) ) { TOL2 = eps2 * fmax( rP, rQ ); }
            if( !( TOL2 > 0.0 ) ) { TOL2 = 1e-300; }
            const double eps = 1.4901161193847656e-08;    // sqrt of the double epsilon
            double v[3] = { cP[0] - cQ[0], cP[1] - cQ[1], cP[2] - cQ[2] };
            double W[4][3];
            int n = 1;
            support3<M>( P, Q, a, b, v, W[0] );
            v[0] = W[0][0]; v[1] = W[0][1]; v[2] = W[0][2];
            double vv = dot3( v, v );
            for( int iter = 0; iter < 64; ++iter )
            {
                if( vv < TOL2 ) { return true; }
                double w[3];
                const double vw = support3<M>( P, Q, a, b, v, w );
                if( vw > 0.0 && vw * vw > TOL2 ) { return false; }
                if( fabs( vv - vw ) <= eps * vv ) { return !( TOL2 < vv ); }
                bool dup = false;
                const double ww = dot3( w, w );
                #pragma unroll
                for( int k = 0; k < 4; ++k )
                {
                    if( k < n )
                    {
                        double d2 = 0;
                        for( int c = 0; c < 3; ++c ) { const double t = W[k][c] - w[c]; d2 += t * t; }
                        dup = dup || ( d2 <= 1e-28 * ( ww + 1e-300 ) );
                    }
                }
                if( dup ) { return !( TOL2 < vv ); }
                #pragma unroll
                for( int c = 0; c < 3; ++c ) { W[3][c] = W[2][c]; W[2][c] = W[1][c]; W[1][c] = W[0][c]; W[0][c] = w[c]; }
                ++n;
                double vn[3];
                if( closest_face( W, n, vn ) ) { return true; }
                const double vvn = dot3( vn, vn );
                v[0] = vn[0]; v[1] = vn[1]; v[2] = vn[2];
                if( fabs( vv - vvn ) <= eps * vvn ) { return !( TOL2 < vvn ); }
                vv = vvn;
            }
            return !( TOL2 < vv );
        }

        // CollisionFinder::FindMaximumSafeStepSize: depth-first bisection of [0, tinit] for the earliest
        // small interval whose hulls meet. The reference keeps the right ends of the intervals still to
        // visit on a stack; here the position is a path of left (0) and right (1) turns, and a pop
        // replays a <- (a + b)/2 or b <- (a + b)/2 from the root, which reproduces the stacked doubles
        // exactly (the same operations on the same operands).
        template<int M>
        __device__ double safe_step_pair( const double * P, const double * Q, double tinit, double eps, unsigned & tests )
        {
            double cp[3] = {0,0,0}, cq[3] = {0,0,0};
            #pragma unroll
            for( int i = 0; i < M; ++i ) { for( int k = 0; k < 3; ++k ) { cp[k] += P[6 * i + k] / M; cq[k] += Q[6 * i + k] / M; } }
            double r2 = 0;
            for( int k = 0; k < 3; ++k ) { r2 += ( cp[k] - cq[k] ) * ( cp[k] - cq[k] ); }
            const double Ts = sqrt( r2 ) / tinit + 1e-300;
            constexpr int MAX_ITER = 126;
            double a = 0.0, b = tinit;
            unsigned long long path0 = 0, path1 = 0;     // turn at depth l: bit l of path0 (l < 64) or path1
            int depth = 0, sp = 0, iter = 0;             // sp: left turns on the path (the reference's stack size)
            while( ( b - a > eps * b ) && ( iter < MAX_ITER ) )
            {
                ++iter; ++tests;
                if( !hulls_meet<M>( P, Q, a, b, Ts ) )
                {
                    if( sp == 0 ) { return b; }
                    int l = depth - 1;                   // deepest left turn
                    while( ( ( l < 64 ? path0 >> l : path1 >> ( l - 64 ) ) & 1ull ) != 0 ) { --l; }
                    if( l < 64 ) { path0 |= 1ull << l; path0 &= ( l == 63 ) ? ~0ull : ( ( 2ull << l ) - 1 ); path1 = 0; }
                    else { path1 |= 1ull << ( l - 64 ); path1 &= ( l == 127 ) ? ~0ull : ( ( 2ull << ( l - 64 ) ) - 1 ); }
                    depth = l + 1; --sp;
                    double A = 0.0, B = tinit;
                    for( int m = 0; m < l; ++m )
                    {
                        const double mid = 0.5 * ( A + B );
                        if( ( ( m < 64 ? path0 >> m : path1 >> ( m - 64 ) ) & 1ull ) != 0 ) { A = mid; } else { B = mid; }
                    }
                    a = b; b = B;
                }
                else
                {
                    ++depth; ++sp;                       // the new turn (bit depth - 1) is left: already 0
                    b = 0.5 * ( a + b );
                }
            }
            if( sp >= MAX_ITER || iter >= MAX_ITER ) { return a; }
            return ( a <= 0.0 ) ? b : a;
        }
    }

    // Per primitive in cluster order: swept box over [0, t_max], vertex ids, moving data (positions, velocities).
    template<int M>
    __global__ __launch_bounds__(256) void k_prim_stage( int64_t T, const int32_t * P_ordering, const int32_t * S, const double * X, const double * U,
                                                         double t_max, double * box, int32_t * SV, double * PU )
    {
        const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
        if( i >= T ) { return; }
        const int64_t t = P_ordering[i];
        double lo[3] = { 1e300, 1e300, 1e300 }, hi[3] = { -1e300, -1e300, -1e300 };
        for( int l = 0; l < M; ++l )
        {
            const int32_t w = S[M * t + l];
            SV[M * i + l] = w;
            for( int k = 0; k < 3; ++k )
            {
                const double x = X[3 * (int64_t)w + k], u = U[3 * (int64_t)w + k];
                const double x1 = x + t_max * u;
                lo[k] = fmin( lo[k], fmin( x, x1 ) ); hi[k] = fmax( hi[k], fmax( x, x1 ) );
                PU[( M * i + l ) * 6 + k] = x; PU[( M * i + l ) * 6 + 3 + k] = u;
            }
        }
        for( int k = 0; k < 3; ++k ) { box[6 * i + k] = lo[k]; box[6 * i + 3 + k] = hi[k]; }
    }

    __global__ __launch_bounds__(256) void k_leaf_boxes( int64_t L, const int32_t * leaf_cluster, const int32_t * C_begin, const int32_t * C_end,
                                                         const double * pbox, double * cbox )
    {
        const int64_t l = (int64_t)blockIdx.x * 256 + threadIdx.x;
        if( l >= L ) { return; }
        const int32_t c = leaf_cluster[l];
        double b[6] = { 1e300, 1e300, 1e300, -1e300, -1e300, -1e300 };
        for( int32_t i = C_begin[c]; i < C_end[c]; ++i )
        {
            for( int k = 0; k < 3; ++k ) { b[k] = fmin( b[k], pbox[6 * (int64_t)i + k] ); b[3 + k] = fmax( b[3 + k], pbox[6 * (int64_t)i + 3 + k] ); }
        }
        for( int k = 0; k < 6; ++k ) { cbox[6 * (int64_t)c + k] = b[k]; }
    }

    __global__ __launch_bounds__(256) void k_level_boxes( int64_t n, const int32_t * nodes, const int32_t * C_left, const int32_t * C_right, double * cbox )
    {
        const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
        if( k >= n ) { return; }
        const int32_t c = nodes[k];
        const int32_t L = C_left[c], R = C_right[c];
        if( L < 0 ) { return; }
        for( int j = 0; j < 3; ++j )
        {
            cbox[6 * (int64_t)c + j] = fmin( cbox[6 * (int64_t)L + j], cbox[6 * (int64_t)R + j] );
            cbox[6 * (int64_t)c + 3 + j] = fmax( cbox[6 * (int64_t)L + 3 + j], cbox[6 * (int64_t)R + 3 + j] );
        }
    }

    // counters: [0] pairs written to `out`, [1] leaf pairs written, [2] overflow flag. Each wave reserves
    // room for all of its lanes' children (and leaf pairs) with one atomic per list.
    __global__ __launch_bounds__(256) void k_traverse_step( int64_t n_in, const int2 * in, int2 * out, int64_t out_cap,
                                                            int2 * leaves, int64_t leaf_cap, unsigned long long * counters,
                                                            const int32_t * C_left, const int32_t * C_right, const double * cbox )
    {
        const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
        const int lane = threadIdx.x & 63;
        int2 kids[4];
        int nk = 0;
        bool leaf = false;
        int2 pr = make_int2( 0, 0 );
        if( k < n_in )
        {
            pr = in[k];
            const int32_t i = pr.x, j = pr.y;
            if( boxes_meet( cbox + 6 * (int64_t)i, cbox + 6 * (int64_t)j ) )
            {
                const int32_t Li = C_left[i], Lj = C_left[j];
                if( Li < 0 && Lj < 0 ) { leaf = true; }
                else if( i == j )
                {
                    const int32_t Ri = C_right[i];
                    kids[0] = make_int2( Li, Li ); kids[1] = make_int2( Li, Ri ); kids[2] = make_int2( Ri, Ri ); nk = 3;
                }
                else if( Li >= 0 && Lj >= 0 )
                {
                    const int32_t Ri = C_right[i], Rj = C_right[j];
                    kids[0] = make_int2( Li, Lj ); kids[1] = make_int2( Li, Rj ); kids[2] = make_int2( Ri, Lj ); kids[3] = make_int2( Ri, Rj ); nk = 4;
                }
                else if( Li >= 0 ) { kids[0] = make_int2( Li, j ); kids[1] = make_int2( C_right[i], j ); nk = 2; }
                else               { kids[0] = make_int2( i, Lj ); kids[1] = make_int2( i, C_right[j] ); nk = 2; }
            }
        }
        // inclusive wave scan of the child counts
        int x = nk;
        #pragma unroll
        for( int off = 1; off < 64; off <<= 1 )
        {
            const int y = __shfl_up( x, off );
            if( lane >= off ) { x += y; }
        }
        const int total = __shfl( x, 63 );
        const unsigned long long lmask = __ballot( leaf );
        unsigned long long kbase = 0, lbase = 0;
        if( lane == 0 )
        {
            if( total ) { kbase = atomicAdd( counters, (unsigned long long)total ); }
            if( lmask ) { lbase = atomicAdd( counters + 1, (unsigned long long)__popcll( lmask ) ); }
        }
        kbase = __shfl( kbase, 0 ); lbase = __shfl( lbase, 0 );
        if( leaf )
        {
            const unsigned long long pos = lbase + __popcll( lmask & ( ( 1ull << lane ) - 1 ) );
            if( pos < (unsigned long long)leaf_cap ) { leaves[pos] = pr; } else { atomicMax( counters + 2, 1ull ); }
        }
        if( nk )
        {
            const unsigned long long pos = kbase + (unsigned long long)( x - nk );
            if( pos + nk <= (unsigned long long)out_cap ) { for( int c = 0; c < nk; ++c ) { out[pos + c] = kids[c]; } }
            else { atomicMax( counters + 2, 1ull ); }
        }
    }

    // Candidate simplex pairs of the leaf pairs. A group of CAND_GROUP lanes takes one leaf pair, a lane one
    // primitive i of the first leaf at a time: if i's swept box meets the second leaf's box (the union of
    // its primitives' boxes) the lane runs over that leaf's primitives j (j > i within one leaf) and keeps
    // the pairs whose boxes meet and that share no vertex. The first sweep counts, the workgroup reserves
    // room with one atomic on `counter` (which counts every candidate, also beyond cap), the second sweep
    // writes the pairs (cluster-ordered primitive indices) to `cand`. Many small groups per wave keep
    // enough independent loads in flight.
    constexpr int CAND_GROUP = 16;

    template<int M, bool COUNT>
    __device__ __forceinline__ void leaf_pair_rows( int2 pr, int32_t ib, int ni, int32_t jb, int nj, int r, const int32_t * SV, const double * pbox,
                                                    const double * cbox, unsigned & count, unsigned long long & pos, int2 * cand, int64_t cap )
    {
        const bool same = pr.x == pr.y;
        for( int a = r; a < ni; a += CAND_GROUP )
        {
            const int32_t i = ib + a;
            double bi[6];
            #pragma unroll
            for( int c = 0; c < 6; ++c ) { bi[c] = pbox[6 * (int64_t)i + c]; }
            if( !same && !boxes_meet( bi, cbox + 6 * (int64_t)pr.y ) ) { continue; }
            int32_t vi[M];
            #pragma unroll
            for( int l = 0; l < M; ++l ) { vi[l] = SV[M * (int64_t)i + l]; }
            for( int b = same ? a + 1 : 0; b < nj; ++b )
            {
                const int32_t j = jb + b;
                if( !boxes_meet( bi, pbox + 6 * (int64_t)j ) ) { continue; }
                bool ok = true;
                #pragma unroll
                for( int r2 = 0; r2 < M; ++r2 )
                {
                    const int32_t y = SV[M * (int64_t)j + r2];
                    #pragma unroll
                    for( int l = 0; l < M; ++l ) { ok = ok && vi[l] != y; }
                }
                if( !ok ) { continue; }
                if( COUNT ) { ++count; }
                else
                {
                    if( pos < (unsigned long long)cap ) { cand[pos] = make_int2( i, j ); }
                    ++pos;
                }
            }
        }
    }

    template<int M>
    __global__ __launch_bounds__(256) void k_leaf_pair_candidates( int64_t n_pairs, const int2 * pairs, const int32_t * C_begin, const int32_t * C_end,
                                                                   const int32_t * SV, const double * pbox, const double * cbox, int2 * cand,
                                                                   int64_t cap, unsigned long long * counter )
    {
        __shared__ unsigned long long s_count[4];
        __shared__ unsigned long long s_base;
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int64_t g = ( (int64_t)blockIdx.x * 256 + threadIdx.x ) / CAND_GROUP;
        const int r = threadIdx.x % CAND_GROUP;
        int2 pr = make_int2( 0, 0 );
        int32_t ib = 0, jb = 0;
        int ni = 0, nj = 0;
        if( g < n_pairs )
        {
            pr = pairs[g];
            ib = C_begin[pr.x]; ni = C_end[pr.x] - ib;
            jb = C_begin[pr.y]; nj = C_end[pr.y] - jb;
        }
        unsigned count = 0;
        unsigned long long pos = 0;
        leaf_pair_rows<M, true>( pr, ib, ni, jb, nj, r, SV, pbox, cbox, count, pos, cand, cap );
        // inclusive wave scan of the counts
        unsigned x = count;
        #pragma unroll
        for( int off = 1; off < 64; off <<= 1 )
        {
            const unsigned y = __shfl_up( x, off );
            if( lane >= off ) { x += y; }
        }
        if( lane == 63 ) { s_count[wave] = x; }
        __syncthreads();
        if( threadIdx.x == 0 )
        {
            const unsigned long long total = s_count[0] + s_count[1] + s_count[2] + s_count[3];
            s_base = total ? atomicAdd( counter, total ) : 0ull;
        }
        __syncthreads();
        if( count == 0 ) { return; }
        pos = s_base + ( x - count );
        for( int w = 0; w < wave; ++w ) { pos += s_count[w]; }
        leaf_pair_rows<M, false>( pr, ib, ni, jb, nj, r, SV, pbox, cbox, count, pos, cand, cap );
    }

    // One lane per candidate pair: the reference's bisection; the minimum by an atomic minimum on the
    // double's bits (non-negative doubles order like their bit patterns), so the result does not depend
    // on the order. result: [0] minimum step (bits), [2] hull tests.
    template<int M>
    __global__ __launch_bounds__(256) void k_pair_steps( int64_t n, const int2 * cand, const double * PU, double t_max, double eps,
                                                         unsigned long long * result )
    {
        const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
        double tmin = t_max;
        unsigned tests = 0;
        if( k < n )
        {
            const int2 c = cand[k];
            tmin = safe_step_pair<M>( PU + 6 * M * (int64_t)c.x, PU + 6 * M * (int64_t)c.y, t_max, eps, tests );
        }
        #pragma unroll
        for( int off = 32; off > 0; off >>= 1 )
        {
            tmin = fmin( tmin, __shfl_xor( tmin, off ) );
            tests += __shfl_xor( tests, off );
        }
        if( ( threadIdx.x & 63 ) == 0 && (int64_t)blockIdx.x * 256 + threadIdx.x < n )
        {
            atomicMin( result, (unsigned long long)__double_as_longlong( tmin ) );
            atomicAdd( result + 2, (unsigned long long)tests );
        }
    }

    // ------------------------------------------------------------------ launchers
    void launch_prim_stage( int n, int64_t T, const int32_t * P_ordering, const int32_t * S, const double * X, const double * U,
                            double t_max, double * box, int32_t * SV, double * PU, hipStream_t s )
    {
        if( T <= 0 ) { return; }
        if( n == 2 ) { hipLaunchKernelGGL( k_prim_stage<3>, col_grid( T ), dim3(256), 0, s, T, P_ordering, S, X, U, t_max, box, SV, PU ); }
        else         { hipLaunchKernelGGL( k_prim_stage<2>, col_grid( T ), dim3(256), 0, s, T, P_ordering, S, X, U, t_max, box, SV, PU ); }
    }

    void launch_leaf_boxes( int64_t L, const int32_t * leaf_cluster, const int32_t * C_begin, const int32_t * C_end, const double * pbox,
                            double * cbox, hipStream_t s )
    {
        if( L > 0 ) { hipLaunchKernelGGL( k_leaf_boxes, col_grid( L ), dim3(256), 0, s, L, leaf_cluster, C_begin, C_end, pbox, cbox ); }
    }

    void launch_level_boxes( int64_t n, const int32_t * nodes, const int32_t * C_left, const int32_t * C_right, double * cbox, hipStream_t s )
    {
        if( n > 0 ) { hipLaunchKernelGGL( k_level_boxes, col_grid( n ), dim3(256), 0, s, n, nodes, C_left, C_right, cbox ); }
    }

    void launch_traverse_step( int64_t n_in, const int2 * in, int2 * out, int64_t out_cap, int2 * leaves, int64_t leaf_cap,
                               unsigned long long * counters, const int32_t * C_left, const int32_t * C_right, const double * cbox, hipStream_t s )
    {
        if( n_in > 0 ) { hipLaunchKernelGGL( k_traverse_step, col_grid( n_in ), dim3(256), 0, s, n_in, in, out, out_cap, leaves, leaf_cap, counters, C_left, C_right, cbox ); }
    }

    void launch_leaf_pair_candidates( int n, int64_t n_pairs, const int2 * pairs, const int32_t * C_begin, const int32_t * C_end,
                                      const int32_t * SV, const double * pbox, const double * cbox, int2 * cand, int64_t cap,
                                      unsigned long long * counter, hipStream_t s )
    {
        if( n_pairs <= 0 ) { return; }
        const dim3 grid( (unsigned)( ( n_pairs * CAND_GROUP + 255 ) / 256 ) );
        if( n == 2 ) { hipLaunchKernelGGL( k_leaf_pair_candidates<3>, grid, dim3(256), 0, s, n_pairs, pairs, C_begin, C_end, SV, pbox, cbox, cand, cap, counter ); }
        else         { hipLaunchKernelGGL( k_leaf_pair_candidates<2>, grid, dim3(256), 0, s, n_pairs, pairs, C_begin, C_end, SV, pbox, cbox, cand, cap, counter ); }
    }

    void launch_pair_steps( int n, int64_t n_cand, const int2 * cand, const double * PU, double t_max, double eps, unsigned long long * result,
                            hipStream_t s )
    {
        if( n_cand <= 0 ) { return; }
        if( n == 2 ) { hipLaunchKernelGGL( k_pair_steps<3>, col_grid( n_cand ), dim3(256), 0, s, n_cand, cand, PU, t_max, eps, result ); }
        else         { hipLaunchKernelGGL( k_pair_steps<2>, col_grid( n_cand ), dim3(256), 0, s, n_cand, cand, PU, t_max, eps, result ); }
    }
}
