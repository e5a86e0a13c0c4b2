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

        // Two moving simplices of M points over [a, b], lifted to R^4 with time scale Ts.
        template<int M>
        struct MovingPair
        {
            double p[M][3], u[M][3], q[M][3], v[M][3];
            double a, b, Ts;
        };

        // argmin over the Minkowski difference D = hull(P) - hull(Q) of d . x
        template<int M>
        __device__ __forceinline__ void support_min( const MovingPair<M> & S, const double d[4], double w[4] )
        {
            double bestP = 0, bestQ = 0, xP[4] = {0,0,0,0}, yQ[4] = {0,0,0,0};
            bool first = true;
            #pragma unroll
            for( int i = 0; i < M; ++i )
            {
                #pragma unroll
                for( int s = 0; s < 2; ++s )
                {
                    const double t = s == 0 ? S.a : S.b;
                    const double x0 = S.p[i][0] + t * S.u[i][0], x1 = S.p[i][1] + t * S.u[i][1], x2 = S.p[i][2] + t * S.u[i][2], x3 = t * S.Ts;
                    const double vp = d[0] * x0 + d[1] * x1 + d[2] * x2 + d[3] * x3;
                    const double y0 = S.q[i][0] + t * S.v[i][0], y1 = S.q[i][1] + t * S.v[i][1], y2 = S.q[i][2] + t * S.v[i][2];
                    const double vq = d[0] * y0 + d[1] * y1 + d[2] * y2 + d[3] * x3;
                    if( first || vp < bestP ) { bestP = vp; xP[0] = x0; xP[1] = x1; xP[2] = x2; xP[3] = x3; }
                    if( first || vq > bestQ ) { bestQ = vq; yQ[0] = y0; yQ[1] = y1; yQ[2] = y2; yQ[3] = x3; }
                    first = false;
                }
            }
            #pragma unroll
            for( int k = 0; k < 4; ++k ) { w[k] = xP[k] - yQ[k]; }
        }

        __device__ __forceinline__ double dot4( const double * a, const double * b ) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3]; }

        // Minimum-norm point of conv(W), |W| = n <= 5 points of R^4: over all subsets, the affine
        // minimum-norm point of the subset if it lies in the subset's relative interior; the smallest of
        // these is the minimum over the hull. W is compacted to the chosen subset.
        __device__ bool min_norm_point( double W[5][4], int & n, double v[4] )
        {
            double best = 1e300;
            int best_mask = 0;
            double bx[4] = {0,0,0,0};
            const int full = ( 1 << n ) - 1;
            for( int mask = 1; mask <= full; ++mask )
            {
                int idx[5], k = 0;
                for( int i = 0; i < n; ++i ) { if( mask & (1 << i) ) { idx[k++] = i; } }
                double lam[5];
                double x[4];
                const double * x0 = W[idx[0]];
                if( k == 1 )
                {
                    lam[0] = 1.0;
                    for( int c = 0; c < 4; ++c ) { x[c] = x0[c]; }
                }
                else
                {
                    double E[4][4], G[4][5];
                    const int r = k - 1;
                    for( int j = 0; j < r; ++j ) { for( int c = 0; c < 4; ++c ) { E[j][c] = W[idx[j + 1]][c] - x0[c]; } }
                    double scale = 0.0;
                    for( int i = 0; i < r; ++i )
                    {
                        for( int j = 0; j < r; ++j ) { G[i][j] = dot4( E[i], E[j] ); }
                        G[i][r] = -dot4( E[i], x0 );
                        scale = fmax( scale, G[i][i] );
                    }
                    bool ok = true;
                    for( int c = 0; c < r && ok; ++c )
                    {
                        int piv = c;
                        for( int i = c + 1; i < r; ++i ) { if( fabs( G[i][c] ) > fabs( G[piv][c] ) ) { piv = i; } }
                        if( !( fabs( G[piv][c] ) > 1e-13 * scale ) ) { ok = false; break; }
                        if( piv != c ) { for( int j = 0; j <= r; ++j ) { const double t = G[c][j]; G[c][j] = G[piv][j]; G[piv][j] = t; } }
                        for( int i = c + 1; i < r; ++i )
                        {
                            const double f = G[i][c] / G[c][c];
                            for( int j = c; j <= r; ++j ) { G[i][j] -= f * G[c][j]; }
                        }
                    }
                    if( !ok ) { continue; }
                    double mu[4];
                    for( int i = r - 1; i >= 0; --i )
                    {
                        double s = G[i][r];
                        for( int j = i + 1; j < r; ++j ) { s -= G[i][j] * mu[j]; }
                        mu[i] = s / G[i][i];
                    }
                    double l0 = 1.0;
                    for( int j = 0; j < r; ++j ) { lam[j + 1] = mu[j]; l0 -= mu[j]; }
                    lam[0] = l0;
                    bool inside = true;
                    for( int j = 0; j < k; ++j ) { inside = inside && lam[j] > 0.0; }
                    if( !inside ) { continue; }
                    for( int c = 0; c < 4; ++c )
                    {
                        double s = x0[c];
                        for( int j = 0; j < r; ++j ) { s += mu[j] * E[j][c]; }
                        x[c] = s;
                    }
                }
                const double n2 = dot4( x, x );
                if( n2 < best ) { best = n2; best_mask = mask; for( int c = 0; c < 4; ++c ) { bx[c] = x[c]; } }
            }
            if( best_mask == 0 ) { return false; }
            int k = 0;
            for( int i = 0; i < n; ++i )
            {
                if( best_mask & (1 << i) ) { if( k != i ) { for( int c = 0; c < 4; ++c ) { W[k][c] = W[i][c]; } } ++k; }
            }
            n = k;
            for( int c = 0; c < 4; ++c ) { v[c] = bx[c]; }
            return true;
        }

        // Do the space-time hulls meet? GJK on D = hull(P) - hull(Q) with the reference's stopping rules:
        // meeting once |v|^2 <= TOL2 or the simplex encloses the origin, separated once a support value
        // exceeds the tolerance on the far side or the progress stalls above it.
        template<int M>
        __device__ bool hulls_meet( const MovingPair<M> & S )
        {
            // interior points and squared radii (tolerance scale)
            const double tm = 0.5 * ( S.a + S.b );
            double cP[4] = {0,0,0,0}, cQ[4] = {0,0,0,0};
            #pragma unroll
            for( int i = 0; i < M; ++i )
            {
                for( int c = 0; c < 3; ++c ) { cP[c] += ( S.p[i][c] + tm * S.u[i][c] ) / M; cQ[c] += ( S.q[i][c] + tm * S.v[i][c] ) / M; }
            }
            cP[3] = cQ[3] = tm * S.Ts;
            double rP = 0, rQ = 0;
            #pragma unroll
            for( int i = 0; i < M; ++i )
            {
                #pragma unroll
                for( int s = 0; s < 2; ++s )
                {
                    const double t = s == 0 ? S.a : S.b;
                    double dp = 0, dq = 0;
                    for( int c = 0; c < 3; ++c )
                    {
                        const double x = S.p[i][c] + t * S.u[i][c] - cP[c]; dp += x * x;
                        const double y = S.q[i][c] + t * S.v[i][c] - cQ[c]; dq += y * y;
                    }
                    const double z = t * S.Ts - cP[3];
                    rP = fmax( rP, dp + z * z ); rQ = fmax( rQ, dq + z * z );
                }
            }
            const double eps2 = 2.220446049250313e-16;
            double TOL2 = eps2 * fmin( rP, rQ );
            if( !( TOL2 > 0.0 ) ) { TOL2 = eps2 * fmax( rP, rQ ); }
            if( !( TOL2 > 0.0 ) ) { TOL2 = 1e-300; }
            double v[4];
            for( int c = 0; c < 4; ++c ) { v[c] = cP[c] - cQ[c]; }
            double W[5][4];
            int n = 0;
            {
                double w[4];
                support_min( S, v, w );
                for( int c = 0; c < 4; ++c ) { W[0][c] = w[c]; v[c] = w[c]; }
                n = 1;
            }
            const double eps = 1.4901161193847656e-08;   // sqrt of the double epsilon
            for( int iter = 0; iter < 64; ++iter )
            {
                const double vv = dot4( v, v );
                if( vv < TOL2 ) { return true; }
                if( n >= 5 ) { return true; }
                double w[4];
                support_min( S, v, w );
                const double vw = dot4( v, w );
                if( vw > 0.0 && vw * vw > TOL2 ) { return false; }
                if( fabs( vv - vw ) <= eps * vv ) { return !( TOL2 < vv ); }
                bool dup = false;
                for( int i = 0; i < n; ++i )
                {
                    double d2 = 0;
                    for( int c = 0; c < 4; ++c ) { const double t = W[i][c] - w[c]; d2 += t * t; }
                    dup = dup || ( d2 <= 1e-28 * ( dot4( w, w ) + 1e-300 ) );
                }
                if( dup ) { return !( TOL2 < vv ); }
                for( int c = 0; c < 4; ++c ) { W[n][c] = w[c]; }
                ++n;
                double vn[4];
                if( !min_norm_point( W, n, vn ) ) { return !( TOL2 < vv ); }
                const double vvn = dot4( vn, vn );
                for( int c = 0; c < 4; ++c ) { v[c] = vn[c]; }
                if( fabs( vv - vvn ) <= eps * vvn ) { return !( TOL2 < vvn ); }
            }
            return !( TOL2 < dot4( v, v ) );
        }

        // CollisionFinder::FindMaximumSafeStepSize
        template<int M>
        __device__ double safe_step_pair( MovingPair<M> & S, double tinit, double eps, int64_t & tests )
        {
            double a = 0.0, b = tinit;
            double cp[3] = {0,0,0}, cq[3] = {0,0,0};
            for( int i = 0; i < M; ++i ) { for( int k = 0; k < 3; ++k ) { cp[k] += S.p[i][k] / M; cq[k] += S.q[i][k] / M; } }
            double r2 = 0;
            for( int k = 0; k < 3; ++k ) { r2 += ( cp[k] - cq[k] ) * ( cp[k] - cq[k] ); }
            S.Ts = sqrt( r2 ) / tinit + 1e-300;
            constexpr int MAX_ITER = 126;
            double stack[MAX_ITER + 2];
            int iter = 0, sp = 0;
            while( ( b - a > eps * b ) && ( iter < MAX_ITER ) )
            {
                ++iter; ++tests;
                S.a = a; S.b = b;
                if( !hulls_meet<M>( S ) )
                {
                    if( sp > 0 ) { a = b; b = stack[sp--]; }
                    else { return b; }
                }
                else { stack[++sp] = b; b = 0.5 * ( a + b ); }
            }
            if( sp >= MAX_ITER || iter >= MAX_ITER ) { return a; }
            return ( a <= 0.0 ) ? b : a;
        }
    }

    __global__ __launch_bounds__(256) void k_prim_swept_boxes( int64_t T, int size, const int32_t * P_ordering, const int32_t * S,
                                                               const double * X, const double * U, double t_max, double * box )
    {
        const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
        if( i >= T ) { return; }
        const int64_t t = P_ordering[i];
        double lo[3] = { 1e300, 1e300, 1e300 }, hi[3] = { -1e300, -1e300, -1e300 };
        for( int l = 0; l < size; ++l )
        {
            const int64_t w = S[size * t + l];
            for( int k = 0; k < 3; ++k )
            {
                const double x0 = X[3 * w + k], x1 = X[3 * w + k] + t_max * U[3 * w + k];
                lo[k] = fmin( lo[k], fmin( x0, x1 ) ); hi[k] = fmax( hi[k], fmax( x0, x1 ) );
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

    // counters: [0] pairs written to `out`, [1] leaf pairs written, [2] overflow flag
    __global__ __launch_bounds__(256) void k_traverse_step( int64_t n_in, const int2 * in, int2 * out, int64_t out_cap,
                                                            int2 * leaves, int64_t leaf_cap, unsigned long long * counters,
                                                            const int32_t * C_left, const int32_t * C_right, const double * cbox )
    {
        const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
        if( k >= n_in ) { return; }
        const int2 pr = in[k];
        const int32_t i = pr.x, j = pr.y;
        if( !boxes_meet( cbox + 6 * (int64_t)i, cbox + 6 * (int64_t)j ) ) { return; }
        const int32_t Li = C_left[i], Lj = C_left[j];
        int2 kids[4];
        int nk = 0;
        if( Li < 0 && Lj < 0 )
        {
            const unsigned long long pos = atomicAdd( counters + 1, 1ull );
            if( pos < (unsigned long long)leaf_cap ) { leaves[pos] = make_int2( i, j ); } else { atomicMax( counters + 2, 1ull ); }
            return;
        }
        if( i == j )
        {
            const int32_t Ri = C_right[i];
            kids[0] = make_int2( Li, Li ); kids[1] = make_int2( Li, Ri ); kids[2] = make_int2( Ri, Ri ); nk = 3;
        }
        else if( Li >= 0 && Lj >= 0 )
        {
            const int32_t Ri = C_right[i], Rj = C_right[j];
            kids[0] = make_int2( Li, Lj ); kids[1] = make_int2( Li, Rj ); kids[2] = make_int2( Ri, Lj ); kids[3] = make_int2( Ri, Rj ); nk = 4;
        }
        else if( Li >= 0 )
        {
            kids[0] = make_int2( Li, j ); kids[1] = make_int2( C_right[i], j ); nk = 2;
        }
        else
        {
            kids[0] = make_int2( i, Lj ); kids[1] = make_int2( i, C_right[j] ); nk = 2;
        }
        const unsigned long long pos = atomicAdd( counters, (unsigned long long)nk );
        if( pos + nk <= (unsigned long long)out_cap ) { for( int c = 0; c < nk; ++c ) { out[pos + c] = kids[c]; } }
        else { atomicMax( counters + 2, 1ull ); }
    }

    // result: [0] minimum step (bits of a non-negative double), [1] candidate pairs, [2] hull tests
    template<int N>
    __global__ __launch_bounds__(256) void k_leaf_pair_steps( int64_t n_pairs, const int2 * pairs, const int32_t * C_begin, const int32_t * C_end,
                                                              const int32_t * P_ordering, const int32_t * S, const double * X, const double * U,
                                                              const double * pbox, double t_max, double eps, unsigned long long * result )
    {
        constexpr int M = N + 1;
        __shared__ double smin[256];
        __shared__ unsigned long long scand[256], stest[256];
        const int64_t b = blockIdx.x;
        double tmin = t_max;
        int64_t cand = 0, tests = 0;
        if( b < n_pairs )
        {
            const int2 pr = pairs[b];
            const int32_t ib = C_begin[pr.x], ie = C_end[pr.x], jb = C_begin[pr.y], je = C_end[pr.y];
            const int ni = ie - ib, nj = je - jb;
            const bool same = pr.x == pr.y;
            for( int k = threadIdx.x; k < ni * nj; k += 256 )
            {
                const int32_t i = ib + k / nj, j = jb + k % nj;
                if( same && j <= i ) { continue; }
                if( !boxes_meet( pbox + 6 * (int64_t)i, pbox + 6 * (int64_t)j ) ) { continue; }
                const int64_t ti = P_ordering[i], tj = P_ordering[j];
                int32_t vi[M], vj[M];
                bool adjacent = false;
                for( int l = 0; l < M; ++l ) { vi[l] = S[M * ti + l]; vj[l] = S[M * tj + l]; }
                for( int l = 0; l < M; ++l ) { for( int r = 0; r < M; ++r ) { adjacent = adjacent || vi[l] == vj[r]; } }
                if( adjacent ) { continue; }
                ++cand;
                MovingPair<M> P;
                for( int l = 0; l < M; ++l )
                {
                    for( int c = 0; c < 3; ++c )
                    {
                        P.p[l][c] = X[3 * (int64_t)vi[l] + c]; P.u[l][c] = U[3 * (int64_t)vi[l] + c];
                        P.q[l][c] = X[3 * (int64_t)vj[l] + c]; P.v[l][c] = U[3 * (int64_t)vj[l] + c];
                    }
                }
                tmin = fmin( tmin, safe_step_pair<M>( P, t_max, eps, tests ) );
            }
        }
        smin[threadIdx.x] = tmin; scand[threadIdx.x] = (unsigned long long)cand; stest[threadIdx.x] = (unsigned long long)tests;
        __syncthreads();
        for( int w = 128; w > 0; w >>= 1 )
        {
            if( (int)threadIdx.x < w )
            {
                smin[threadIdx.x] = fmin( smin[threadIdx.x], smin[threadIdx.x + w] );
                scand[threadIdx.x] += scand[threadIdx.x + w];
                stest[threadIdx.x] += stest[threadIdx.x + w];
            }
            __syncthreads();
        }
        if( threadIdx.x == 0 && b < n_pairs )
        {
            atomicMin( result, (unsigned long long)__double_as_longlong( smin[0] ) );
            atomicAdd( result + 1, scand[0] );
            atomicAdd( result + 2, stest[0] );
        }
    }

    // ------------------------------------------------------------------ launchers
    void launch_prim_swept_boxes( int64_t T, int size, const int32_t * P_ordering, const int32_t * S, const double * X, const double * U,
                                  double t_max, double * box, hipStream_t s )
    {
        if( T > 0 ) { hipLaunchKernelGGL( k_prim_swept_boxes, col_grid( T ), dim3(256), 0, s, T, size, P_ordering, S, X, U, t_max, box ); }
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

    void launch_leaf_pair_steps( int n, int64_t n_pairs, const int2 * pairs, const int32_t * C_begin, const int32_t * C_end,
                                 const int32_t * P_ordering, const int32_t * S, const double * X, const double * U, const double * pbox,
                                 double t_max, double eps, unsigned long long * result, hipStream_t s )
    {
        if( n_pairs <= 0 ) { return; }
        const dim3 grid( (unsigned)n_pairs );
        if( n == 2 ) { hipLaunchKernelGGL( k_leaf_pair_steps<2>, grid, dim3(256), 0, s, n_pairs, pairs, C_begin, C_end, P_ordering, S, X, U, pbox, t_max, eps, result ); }
        else         { hipLaunchKernelGGL( k_leaf_pair_steps<1>, grid, dim3(256), 0, s, n_pairs, pairs, C_begin, C_end, P_ordering, S, X, U, pbox, t_max, eps, result ); }
    }
}
