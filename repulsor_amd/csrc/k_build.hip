// Device rebuild of the cluster tree and the symmetric block cluster tree from the vertex coordinates.
//
//   ClusterTree construction (split + depth-first serialisation)
//       src/ClusterTree.hpp:115-180, src/ClusterTree/Split_Thread.hpp:3-166, src/ClusterTree/Serialize_Thread.hpp:3-156
//   AABB_LongestAxisSplit::split          src/BoundingVolumes/AABB_LongestAxisSplit.hpp:134-253
//   AABB::FromPrimitives                  src/BoundingVolumes/AABB.hpp:155-237
//   BlockClusterTree::ComputeBlocks       src/BlockClusterTree.hpp:234-330
//   Traverse_DepthFirst, BlockSplit_Kernel src/ClusterTreePairTraversor/Traverse_DepthFirst.hpp:3-267,
//                                          src/BlockClusterTree/BlockSplit_Kernel.hpp:164-251
//   AABB MAC                               src/GJK.hpp:1175-1183, src/BoundingVolumes/AABB.hpp:407-430
//
// The outputs equal the host restatement's (include/repulsor/ClusterTree.hpp, BlockClusterTree.hpp)
// element for element: the arithmetic follows it operation by operation with floating-point
// contraction off, and every order-dependent step is restated exactly:
//
// * Tree, one depth level per round. A node's box is the min/max over its primitives' hull points
//   (order-free). The split is the reference's in-place partition: scanning the range, an element whose
//   interior point lies below the box center along the longest axis is swapped to the front. Its result
//   is a permutation that the kernels compute in parallel: a "below" element at scan index k goes to
//   begin + (number of "below" elements before k); the other elements form a FIFO queue that every later
//   "below" element rotates by one, so the element that ends in the r-th slot of the right part is found
//   by following tape[t] = tape[t - g_k] (g_k = "above" elements before k) back to an "above" element.
//   When every interior point lies on one side, the median fallback sorts the range stably by the
//   interior coordinate (bitonic sort on (coordinate, position) in LDS) and splits at the middle.
// * Depth-first numbering from descendant counts (children follow their parent, the left subtree first).
// * Block cluster tree, one traversal level per launch over cluster pairs (the reference's
//   depth-first order only changes the order of the lists, which are sorted afterwards): admissible
//   pairs become far blocks, leaf-leaf pairs leaf blocks; the primitive pairs of every leaf block are
//   classified (adjacent or separated by the primitive MAC: near; else very near or intersecting) with
//   the host's exact simplex distance restated below. Lists are sorted with radix sorts (hipCUB).
// * Near tiles and the far CSR are assembled from the sorted lists.

#include "rp_launch.hpp"
#include "rp_capi_util.hpp"

#include <hipcub/hipcub.hpp>

#include <cfloat>
#include <climits>

namespace rp
{
    namespace
    {
        constexpr int BT_WG = 256;
        constexpr int BT_SORT_MAX = 2048;   // largest range the median fallback sorts in LDS

        inline dim3 bt_grid( int64_t n ) { return dim3( (unsigned)( ( n + BT_WG - 1 ) / BT_WG ) ); }

        // std::min( a, x ) and std::max( a, x ) of the host (the first argument wins ties)
        __device__ __forceinline__ double bt_min( double a, double x ) { return ( x < a ) ? x : a; }
        __device__ __forceinline__ double bt_max( double a, double x ) { return ( a < x ) ? x : a; }

        // ---------------------------------------------------------------- primitive data
        // Serialized primitive [r^2, interior point (3), hull (S x 3)] in the caller's order (Polytope
        // FromIndexList), and the simplex with its vertices sorted ascending.
        template<int S>
        __global__ __launch_bounds__(BT_WG) void k_bt_serialize( int64_t T, const double * __restrict__ V, const int32_t * __restrict__ simp,
                                                                 double * __restrict__ pser, int32_t * __restrict__ ssimp )
        {
            _Pragma("clang fp contract(off)")
            constexpr int PS = 4 + 3 * S;
            const int64_t i = (int64_t)blockIdx.x * BT_WG + threadIdx.x;
            if( i >= T ) { return; }
            int32_t v[S];
            double h[S][3];
            #pragma unroll
            for( int j = 0; j < S; ++j )
            {
                v[j] = simp[S * i + j];
                #pragma unroll
                for( int k = 0; k < 3; ++k ) { h[j][k] = V[3 * (int64_t)v[j] + k]; }
            }
            const double w = 1.0 / S;
            double c[3];
            #pragma unroll
            for( int k = 0; k < 3; ++k )
            {
                c[k] = h[0][k];
                #pragma unroll
                for( int j = 1; j < S; ++j ) { c[k] += h[j][k]; }
                c[k] *= w;
            }
            double r2 = 0.0;
            #pragma unroll
            for( int j = 0; j < S; ++j )
            {
                double d = h[j][0] - c[0];
                double sq = d * d;
                #pragma unroll
                for( int k = 1; k < 3; ++k ) { d = h[j][k] - c[k]; sq += d * d; }
                r2 = bt_max( r2, sq );
            }
            double * p = pser + PS * i;
            p[0] = r2;
            #pragma unroll
            for( int k = 0; k < 3; ++k ) { p[1 + k] = c[k]; }
            #pragma unroll
            for( int j = 0; j < S; ++j )
            {
                #pragma unroll
                for( int k = 0; k < 3; ++k ) { p[4 + 3 * j + k] = h[j][k]; }
            }
            #pragma unroll
            for( int a = 1; a < S; ++a )
            {
                #pragma unroll
                for( int b = a; b > 0; --b )
                {
                    if( v[b] < v[b - 1] ) { const int32_t t = v[b]; v[b] = v[b - 1]; v[b - 1] = t; }
                }
            }
            #pragma unroll
            for( int j = 0; j < S; ++j ) { ssimp[S * i + j] = v[j]; }
        }

        __global__ __launch_bounds__(BT_WG) void k_bt_iota( int64_t T, int32_t * ord, int32_t * pnode )
        {
            const int64_t i = (int64_t)blockIdx.x * BT_WG + threadIdx.x;
            if( i < T ) { ord[i] = (int32_t)i; pnode[i] = 0; }
        }

        // ---------------------------------------------------------------- tree levels
        struct LevelArgs
        {
            int64_t lvl0;          // first node of the level
            int32_t * nb, * ne;    // node ranges
            int32_t * ndepth;
            int32_t * nleft, * nright;
            int8_t  * nflag;       // 0 leaf, 1 partition split, 2 median fallback
            int8_t  * ndir;
            double  * nmid;
            double  * nbox;        // 7 per node
            int32_t * nnl;         // left size
            int32_t * nj0;         // first "above" element, relative to the node's begin
            int32_t * lvl_ind;     // per level node: 1 if it splits
            int32_t * lvl_incl;    // inclusive prefix of lvl_ind
            const int32_t * ord;
            int32_t * ord_new;
            const int32_t * pnode;
            int32_t * pnode_new;
            int32_t * gcnt;        // per position: "above" elements before it in its node
            uint8_t * isl;         // per position: below the split value
            int32_t * fb_list;     // fallback nodes of the level
            unsigned int * fb_ctr; // [0] count, [1] largest range
            const double * pser;
            int thr;
        };

        // AABB::FromPrimitives over the hull points of the node's primitives, then the split decision.
        template<int S>
        __global__ __launch_bounds__(BT_WG) void k_bt_node_box( LevelArgs A )
        {
            _Pragma("clang fp contract(off)")
            constexpr int PS = 4 + 3 * S;
            const int64_t node = A.lvl0 + blockIdx.x;
            const int32_t b = A.nb[node], e = A.ne[node];
            double lo[3], hi[3];
            #pragma unroll
            for( int k = 0; k < 3; ++k ) { lo[k] = DBL_MAX; hi[k] = -DBL_MAX; }
            for( int32_t i = b + (int32_t)threadIdx.x; i < e; i += BT_WG )
            {
                const double * p = A.pser + PS * (int64_t)A.ord[i] + 4;
                #pragma unroll
                for( int j = 0; j < S; ++j )
                {
                    #pragma unroll
                    for( int k = 0; k < 3; ++k ) { const double x = p[3 * j + k]; lo[k] = bt_min( lo[k], x ); hi[k] = bt_max( hi[k], x ); }
                }
            }
            #pragma unroll
            for( int off = 32; off > 0; off >>= 1 )
            {
                #pragma unroll
                for( int k = 0; k < 3; ++k )
                {
                    lo[k] = bt_min( lo[k], __shfl_xor( lo[k], off ) );
                    hi[k] = bt_max( hi[k], __shfl_xor( hi[k], off ) );
                }
            }
            __shared__ double red[BT_WG / 64][6];
            const int wave = threadIdx.x >> 6;
            if( ( threadIdx.x & 63 ) == 0 )
            {
                #pragma unroll
                for( int k = 0; k < 3; ++k ) { red[wave][k] = lo[k]; red[wave][3 + k] = hi[k]; }
            }
            __syncthreads();
            if( threadIdx.x != 0 ) { return; }
            for( int w = 1; w < BT_WG / 64; ++w )
            {
                #pragma unroll
                for( int k = 0; k < 3; ++k ) { lo[k] = bt_min( lo[k], red[w][k] ); hi[k] = bt_max( hi[k], red[w][3 + k] ); }
            }
            double r2 = 0.0;
            #pragma unroll
            for( int k = 0; k < 3; ++k )
            {
                const double diff = 0.5 * ( hi[k] - lo[k] );
                r2 += diff * diff;
                lo[k] += diff;
                hi[k] = diff;
            }
            double * bx = A.nbox + 7 * node;
            bx[0] = r2;
            #pragma unroll
            for( int k = 0; k < 3; ++k ) { bx[1 + k] = lo[k]; bx[4 + k] = hi[k]; }
            const bool split = ( e - b ) > A.thr;
            A.nflag[node] = split ? 1 : 0;
            A.lvl_ind[blockIdx.x] = split ? 1 : 0;
            if( split )
            {
                int dir = 0;
                double Lmax = hi[0];
                #pragma unroll
                for( int k = 1; k < 3; ++k ) { if( hi[k] > Lmax ) { Lmax = hi[k]; dir = k; } }
                A.ndir[node] = (int8_t)dir;
                A.nmid[node] = lo[dir];
            }
        }

        // Partition scan of a splitting node: per position the "below" flag and the number of "above"
        // elements before it; the left size and the first "above" position. Ranges with everything on one
        // side are handed to the median fallback.
        template<int S>
        __global__ __launch_bounds__(BT_WG) void k_bt_node_scan( LevelArgs A )
        {
            constexpr int PS = 4 + 3 * S;
            const int64_t node = A.lvl0 + blockIdx.x;
            if( A.nflag[node] != 1 ) { return; }
            const int32_t b = A.nb[node], e = A.ne[node];
            const int dir = A.ndir[node];
            const double mid = A.nmid[node];
            const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
            __shared__ int wsum[BT_WG / 64];
            __shared__ int j0s;
            if( threadIdx.x == 0 ) { j0s = -1; }
            __syncthreads();
            int carry = 0;
            for( int32_t base = b; base < e; base += BT_WG )
            {
                const int32_t k = base + (int32_t)threadIdx.x;
                const bool valid = k < e;
                bool below = false;
                if( valid ) { below = A.pser[PS * (int64_t)A.ord[k] + 1 + dir] < mid; }
                const bool above = valid && !below;
                const unsigned long long m = __ballot( above );
                const int lp = __popcll( m & ( ( 1ull << lane ) - 1ull ) );
                if( lane == 0 ) { wsum[wave] = __popcll( m ); }
                __syncthreads();
                int woff = 0, tot = 0;
                #pragma unroll
                for( int w = 0; w < BT_WG / 64; ++w ) { woff += ( w < wave ) ? wsum[w] : 0; tot += wsum[w]; }
                const int g = carry + woff + lp;
                if( valid ) { A.gcnt[k] = g; A.isl[k] = below ? 1 : 0; }
                if( above && g == 0 ) { j0s = k - b; }
                carry += tot;
                __syncthreads();
            }
            if( threadIdx.x != 0 ) { return; }
            const int32_t n = e - b;
            const int32_t nl = n - carry;
            if( nl == 0 || nl == n )
            {
                A.nflag[node] = 2;
                A.nnl[node] = n / 2;
                const unsigned slot = atomicAdd( A.fb_ctr, 1u );
                A.fb_list[slot] = (int32_t)node;
                atomicMax( A.fb_ctr + 1, (unsigned)n );
            }
            else
            {
                A.nnl[node] = nl;
                A.nj0[node] = j0s;
            }
        }

        // Children of the splitting nodes, in level order of their parents.
        __global__ __launch_bounds__(BT_WG) void k_bt_children( LevelArgs A, int64_t n_lvl, int64_t next0 )
        {
            const int64_t i = (int64_t)blockIdx.x * BT_WG + threadIdx.x;
            if( i >= n_lvl ) { return; }
            const int64_t node = A.lvl0 + i;
            if( A.nflag[node] == 0 ) { A.nleft[node] = -1; A.nright[node] = -1; return; }
            const int64_t c = next0 + 2 * (int64_t)( A.lvl_incl[i] - 1 );
            const int32_t b = A.nb[node], e = A.ne[node], s = b + A.nnl[node];
            A.nb[c] = b; A.ne[c] = s; A.nb[c + 1] = s; A.ne[c + 1] = e;
            A.ndepth[c] = A.ndepth[c + 1] = A.ndepth[node] + 1;
            A.nleft[node] = (int32_t)c; A.nright[node] = (int32_t)( c + 1 );
        }

        // New order and node of every position.
        __global__ __launch_bounds__(BT_WG) void k_bt_permute( LevelArgs A, int64_t T )
        {
            const int64_t k = (int64_t)blockIdx.x * BT_WG + threadIdx.x;
            if( k >= T ) { return; }
            const int32_t node = A.pnode[k];
            const int f = ( node >= 0 ) ? A.nflag[node] : 0;
            if( f == 0 ) { A.ord_new[k] = A.ord[k]; A.pnode_new[k] = -1; return; }
            const int32_t b = A.nb[node], nl = A.nnl[node];
            const int32_t rel = (int32_t)k - b;
            A.pnode_new[k] = ( rel < nl ) ? A.nleft[node] : A.nright[node];
            if( f == 2 ) { return; }
            if( A.isl[k] ) { A.ord_new[b + rel - A.gcnt[k]] = A.ord[k]; }
            if( rel >= nl )
            {
                // slot rel - nl of the right part holds tape entry (n - j0 - m) + (rel - nl), m = n - nl
                const int32_t j0 = A.nj0[node];
                int32_t t = ( nl - j0 ) + ( rel - nl );
                int32_t kk;
                while( true )
                {
                    kk = b + j0 + t;
                    if( !A.isl[kk] ) { break; }
                    t -= A.gcnt[kk];
                }
                A.ord_new[k] = A.ord[kk];
            }
        }

        // Median fallback: stable sort of the range by the interior coordinate along the split axis.
        template<int S>
        __global__ __launch_bounds__(BT_WG) void k_bt_fallback_sort( LevelArgs A )
        {
            constexpr int PS = 4 + 3 * S;
            const int32_t node = A.fb_list[blockIdx.x];
            const int32_t b = A.nb[node], n = A.ne[node] - b;
            const int dir = A.ndir[node];
            __shared__ double key[BT_SORT_MAX];
            __shared__ int32_t idx[BT_SORT_MAX];
            int P2 = 1;
            while( P2 < n ) { P2 <<= 1; }
            for( int i = threadIdx.x; i < P2; i += BT_WG )
            {
                key[i] = ( i < n ) ? A.pser[PS * (int64_t)A.ord[b + i] + 1 + dir] : HUGE_VAL;
                idx[i] = i;
            }
            __syncthreads();
            for( int k = 2; k <= P2; k <<= 1 )
            {
                for( int j = k >> 1; j > 0; j >>= 1 )
                {
                    for( int i = threadIdx.x; i < P2; i += BT_WG )
                    {
                        const int l = i ^ j;
                        if( l > i )
                        {
                            const double ki = key[i], kl = key[l];
                            const int32_t ii = idx[i], il = idx[l];
                            const bool greater = ( ki > kl ) || ( ki == kl && ii > il );
                            if( ( ( i & k ) == 0 ) == greater )
                            {
                                key[i] = kl; key[l] = ki; idx[i] = il; idx[l] = ii;
                            }
                        }
                    }
                    __syncthreads();
                }
            }
            for( int i = threadIdx.x; i < n; i += BT_WG ) { A.ord_new[b + i] = A.ord[b + idx[i]]; }
        }

        // ---------------------------------------------------------------- depth-first numbering
        __global__ __launch_bounds__(BT_WG) void k_bt_desc( int64_t lvl0, int64_t n, const int8_t * nflag, const int32_t * nleft, const int32_t * nright,
                                                            int32_t * desc, int32_t * dleaf )
        {
            const int64_t i = (int64_t)blockIdx.x * BT_WG + threadIdx.x;
            if( i >= n ) { return; }
            const int64_t node = lvl0 + i;
            if( nflag[node] == 0 ) { desc[node] = 1; dleaf[node] = 1; return; }
            const int32_t l = nleft[node], r = nright[node];
            desc[node] = 1 + desc[l] + desc[r];
            dleaf[node] = dleaf[l] + dleaf[r];
        }

        __global__ __launch_bounds__(BT_WG) void k_bt_dfs( int64_t lvl0, int64_t n, const int8_t * nflag, const int32_t * nleft, const int32_t * nright,
                                                           const int32_t * desc, const int32_t * dleaf, int32_t * g, int32_t * lb )
        {
            const int64_t i = (int64_t)blockIdx.x * BT_WG + threadIdx.x;
            if( i >= n ) { return; }
            const int64_t node = lvl0 + i;
            if( nflag[node] == 0 ) { return; }
            const int32_t l = nleft[node], r = nright[node];
            g[l] = g[node] + 1;
            g[r] = g[node] + 1 + desc[l];
            lb[l] = lb[node];
            lb[r] = lb[node] + dleaf[l];
        }

        struct TreeOut
        {
            int32_t * C_left, * C_right, * C_begin, * C_end, * C_depth;
            double  * C_ser;          // 7 per cluster
            int32_t * leaf_cluster;   // leaf -> cluster
            int32_t * leaf_lookup;    // cluster -> leaf or -1
            int32_t * leaf_b, * leaf_e;
        };

        __global__ __launch_bounds__(BT_WG) void k_bt_scatter( int64_t N, const int32_t * nb, const int32_t * ne, const int32_t * ndepth, const int8_t * nflag,
                                                               const int32_t * nleft, const int32_t * nright, const double * nbox,
                                                               const int32_t * g, const int32_t * lb, TreeOut O )
        {
            const int64_t node = (int64_t)blockIdx.x * BT_WG + threadIdx.x;
            if( node >= N ) { return; }
            const int32_t c = g[node];
            O.C_begin[c] = nb[node]; O.C_end[c] = ne[node]; O.C_depth[c] = ndepth[node];
            #pragma unroll
            for( int k = 0; k < 7; ++k ) { O.C_ser[7 * (int64_t)c + k] = nbox[7 * node + k]; }
            if( nflag[node] == 0 )
            {
                O.C_left[c] = -1; O.C_right[c] = -1;
                const int32_t l = lb[node];
                O.leaf_cluster[l] = c; O.leaf_lookup[c] = l;
                O.leaf_b[l] = nb[node]; O.leaf_e[l] = ne[node];
            }
            else
            {
                O.C_left[c] = g[nleft[node]]; O.C_right[c] = g[nright[node]];
                O.leaf_lookup[c] = -1;
            }
        }

        // Tree-ordered serialized primitives and sorted simplices (the pre-operator columns).
        template<int S>
        __global__ __launch_bounds__(BT_WG) void k_bt_tree_order( int64_t T, const int32_t * ord, const double * pser, const int32_t * ssimp,
                                                                  double * tser, int32_t * pre_cols )
        {
            constexpr int PS = 4 + 3 * S;
            const int64_t i = (int64_t)blockIdx.x * BT_WG + threadIdx.x;
            if( i >= T ) { return; }
            const int64_t o = ord[i];
            #pragma unroll
            for( int k = 0; k < PS; ++k ) { tser[PS * i + k] = pser[PS * o + k]; }
            #pragma unroll
            for( int j = 0; j < S; ++j ) { pre_cols[S * i + j] = ssimp[S * o + j]; }
        }

        // ---------------------------------------------------------------- block cluster tree traversal
        // AdmissableQ: max(r_S^2, r_T^2) < theta^2 dist(AABB_S, AABB_T)^2.
        __device__ __forceinline__ bool bt_admissable( const double * P, const double * Q, double far_theta2 )
        {
            _Pragma("clang fp contract(off)")
            double d2 = 0.0;
            #pragma unroll
            for( int k = 0; k < 3; ++k )
            {
                const double lo = bt_max( P[1 + k] - P[4 + k], Q[1 + k] - Q[4 + k] );
                const double hi = bt_min( P[1 + k] + P[4 + k], Q[1 + k] + Q[4 + k] );
                const double x = bt_max( 0.0, lo - hi );
                d2 += x * x;
            }
            return bt_max( P[0], Q[0] ) < far_theta2 * d2;
        }

        struct TravArgs
        {
            const int2 * in;
            int64_t n_in;
            int2 * out;
            int64_t out_cap;
            int2 * far;
            int64_t far_cap;
            int2 * blocks;
            int64_t blk_cap;
            unsigned long long * ctr;   // [0] next level, [1] far, [2] blocks, [3] overflow
            const int32_t * C_left, * C_right, * leaf_lookup;
            const double * C_ser;
            double far_theta2;
        };

        __device__ __forceinline__ unsigned long long bt_wave_reserve( int cnt, unsigned long long * ctr, int lane, int & excl )
        {
            int x = cnt;
            #pragma unroll
            for( int off = 1; off < 64; off <<= 1 )
            {
                const int y = __shfl_up( x, off );
                if( lane >= off ) { x += y; }
            }
            const int total = __shfl( x, 63 );
            unsigned long long base = 0;
            if( lane == 0 && total ) { base = atomicAdd( ctr, (unsigned long long)total ); }
            excl = x - cnt;
            return __shfl( base, 0 );
        }

        __global__ __launch_bounds__(BT_WG) void k_bt_traverse( TravArgs A )
        {
            const int64_t k = (int64_t)blockIdx.x * BT_WG + threadIdx.x;
            const int lane = threadIdx.x & 63;
            int2 kids[4];
            int nk = 0;
            int nfar = 0, nblk = 0;
            int2 farp = make_int2( 0, 0 ), blk = make_int2( 0, 0 );
            if( k < A.n_in )
            {
                const int2 pr = A.in[k];
                const int32_t i = pr.x, j = pr.y;
                if( i != j && bt_admissable( A.C_ser + 7 * (int64_t)i, A.C_ser + 7 * (int64_t)j, A.far_theta2 ) )
                {
                    nfar = 1; farp = make_int2( min( i, j ), max( i, j ) );
                }
                else
                {
                    const int32_t Li = A.C_left[i], Lj = A.C_left[j];
                    if( Li >= 0 || Lj >= 0 )
                    {
                        const int32_t Ri = A.C_right[i], Rj = A.C_right[j];
                        const double si = ( Li >= 0 ) ? A.C_ser[7 * (int64_t)i] : 0.0;
                        const double sj = ( Lj >= 0 ) ? A.C_ser[7 * (int64_t)j] : 0.0;
                        if( si == sj && si > 0.0 )
                        {
                            if( i == j ) { kids[0] = make_int2( Li, Rj ); kids[1] = make_int2( Ri, Rj ); kids[2] = make_int2( Li, Lj ); nk = 3; }
                            else
                            {
                                kids[0] = make_int2( Ri, Rj ); kids[1] = make_int2( Li, Rj ); kids[2] = make_int2( Ri, Lj ); kids[3] = make_int2( Li, Lj ); nk = 4;
                            }
                        }
                        else if( si > sj ) { kids[0] = make_int2( Ri, j ); kids[1] = make_int2( Li, j ); nk = 2; }
                        else               { kids[0] = make_int2( i, Rj ); kids[1] = make_int2( i, Lj ); nk = 2; }
                    }
                    else
                    {
                        const int32_t li = A.leaf_lookup[i], lj = A.leaf_lookup[j];
                        nblk = 1; blk = make_int2( min( li, lj ), max( li, lj ) );
                    }
                }
            }
            int ek, ef, eb;
            const unsigned long long bk = bt_wave_reserve( nk, A.ctr + 0, lane, ek );
            const unsigned long long bf = bt_wave_reserve( nfar, A.ctr + 1, lane, ef );
            const unsigned long long bb = bt_wave_reserve( nblk, A.ctr + 2, lane, eb );
            if( nk )
            {
                const unsigned long long pos = bk + (unsigned long long)ek;
                if( pos + nk <= (unsigned long long)A.out_cap ) { for( int c = 0; c < nk; ++c ) { A.out[pos + c] = kids[c]; } }
                else { atomicMax( A.ctr + 3, 1ull ); }
            }
            if( nfar )
            {
                const unsigned long long pos = bf + (unsigned long long)ef;
                if( pos < (unsigned long long)A.far_cap ) { A.far[pos] = farp; } else { atomicMax( A.ctr + 3, 1ull ); }
            }
            if( nblk )
            {
                const unsigned long long pos = bb + (unsigned long long)eb;
                if( pos < (unsigned long long)A.blk_cap ) { A.blocks[pos] = blk; } else { atomicMax( A.ctr + 3, 1ull ); }
            }
        }

        __global__ __launch_bounds__(BT_WG) void k_bt_pair_keys( int64_t n, const int2 * p, uint64_t * keys )
        {
            const int64_t i = (int64_t)blockIdx.x * BT_WG + threadIdx.x;
            if( i < n ) { keys[i] = ( (uint64_t)(uint32_t)p[i].x << 32 ) | (uint64_t)(uint32_t)p[i].y; }
        }

        // ---------------------------------------------------------------- primitive pair classification
        struct D3 { double x, y, z; };
        __device__ __forceinline__ D3 d3sub( D3 a, D3 b ) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
        __device__ __forceinline__ D3 d3add( D3 a, D3 b ) { return { a.x + b.x, a.y + b.y, a.z + b.z }; }
        __device__ __forceinline__ D3 d3mul( double s, D3 a ) { return { s * a.x, s * a.y, s * a.z }; }
        __device__ __forceinline__ double d3dot( D3 a, D3 b ) { return a.x * b.x + a.y * b.y + a.z * b.z; }
        __device__ __forceinline__ D3 d3cross( D3 a, D3 b ) { return { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x }; }
        __device__ __forceinline__ double d3clamp01( double t ) { return t < 0 ? 0 : ( t > 1 ? 1 : t ); }

        // Closest point on triangle abc to p (Ericson 5.1.5), the host's expression order.
        __device__ D3 bt_closest_on_triangle( D3 p, D3 a, D3 b, D3 c )
        {
            _Pragma("clang fp contract(off)")
            const D3 ab = d3sub( b, a ), ac = d3sub( c, a ), ap = d3sub( p, a );
            const double d1 = d3dot( ab, ap ), d2 = d3dot( ac, ap );
            if( d1 <= 0 && d2 <= 0 ) { return a; }
            const D3 bp = d3sub( p, b );
            const double d3 = d3dot( ab, bp ), d4 = d3dot( ac, bp );
            if( d3 >= 0 && d4 <= d3 ) { return b; }
            const double vc = d1 * d4 - d3 * d2;
            if( vc <= 0 && d1 >= 0 && d3 <= 0 ) { const double v = d1 / ( d1 - d3 ); return d3add( a, d3mul( v, ab ) ); }
            const D3 cp = d3sub( p, c );
            const double d5 = d3dot( ab, cp ), d6 = d3dot( ac, cp );
            if( d6 >= 0 && d5 <= d6 ) { return c; }
            const double vb = d5 * d2 - d1 * d6;
            if( vb <= 0 && d2 >= 0 && d6 <= 0 ) { const double w = d2 / ( d2 - d6 ); return d3add( a, d3mul( w, ac ) ); }
            const double va = d3 * d6 - d5 * d4;
            if( va <= 0 && ( d4 - d3 ) >= 0 && ( d5 - d6 ) >= 0 )
            {
                const double w = ( d4 - d3 ) / ( ( d4 - d3 ) + ( d5 - d6 ) );
                return d3add( b, d3mul( w, d3sub( c, b ) ) );
            }
            const double denom = 1.0 / ( va + vb + vc );
            const double v = vb * denom, w = vc * denom;
            return d3add( d3add( a, d3mul( v, ab ) ), d3mul( w, ac ) );
        }

        // Squared distance of the segments [p1,q1] and [p2,q2] (Ericson 5.1.9), the host's expression order.
        __device__ double bt_seg_seg_d2( D3 p1, D3 q1, D3 p2, D3 q2 )
        {
            _Pragma("clang fp contract(off)")
            const D3 d1 = d3sub( q1, p1 ), d2 = d3sub( q2, p2 ), r = d3sub( p1, p2 );
            const double a = d3dot( d1, d1 ), e = d3dot( d2, d2 ), f = d3dot( d2, r );
            double s, t;
            const double EPS = 1e-300;
            if( a <= EPS && e <= EPS ) { return d3dot( r, r ); }
            if( a <= EPS ) { s = 0; t = d3clamp01( f / e ); }
            else
            {
                const double c = d3dot( d1, r );
                if( e <= EPS ) { t = 0; s = d3clamp01( -c / a ); }
                else
                {
                    const double b = d3dot( d1, d2 ), denom = a * e - b * b;
                    s = ( denom != 0 ) ? d3clamp01( ( b * f - c * e ) / denom ) : 0.0;
                    t = ( b * s + f ) / e;
                    if( t < 0 ) { t = 0; s = d3clamp01( -c / a ); }
                    else if( t > 1 ) { t = 1; s = d3clamp01( ( b - c ) / a ); }
                }
            }
            const D3 c1 = d3add( p1, d3mul( s, d1 ) ), c2 = d3add( p2, d3mul( t, d2 ) );
            const D3 dd = d3sub( c1, c2 );
            return d3dot( dd, dd );
        }

        // Does the segment [p,q] cross the triangle (a,b,c)? (Moller-Trumbore on the segment.)
        __device__ bool bt_segment_crosses_triangle( D3 p, D3 q, D3 a, D3 b, D3 c )
        {
            _Pragma("clang fp contract(off)")
            const D3 dir = d3sub( q, p ), e1 = d3sub( b, a ), e2 = d3sub( c, a );
            const D3 h = d3cross( dir, e2 );
            const double det = d3dot( e1, h );
            const double scale = sqrt( d3dot( e1, e1 ) * d3dot( dir, dir ) * d3dot( e2, e2 ) );
            if( fabs( det ) <= 1e-14 * scale ) { return false; }
            const double inv = 1.0 / det;
            const D3 s = d3sub( p, a );
            const double u = inv * d3dot( s, h );
            if( u < 0 || u > 1 ) { return false; }
            const D3 qq = d3cross( s, e1 );
            const double v = inv * d3dot( dir, qq );
            if( v < 0 || u + v > 1 ) { return false; }
            const double t = inv * d3dot( e2, qq );
            return ( t >= 0 && t <= 1 );
        }

        // Exact squared distance of two simplices with S points each (SimplexSquaredDistance of the host).
        template<int S>
        __device__ double bt_simplex_d2( const D3 * p, const D3 * q )
        {
            _Pragma("clang fp contract(off)")
            if constexpr ( S == 3 )
            {
                #pragma unroll
                for( int i = 0; i < 3; ++i )
                {
                    if( bt_segment_crosses_triangle( p[i], p[(i + 1) % 3], q[0], q[1], q[2] ) ) { return 0.0; }
                    if( bt_segment_crosses_triangle( q[i], q[(i + 1) % 3], p[0], p[1], p[2] ) ) { return 0.0; }
                }
            }
            double d2 = DBL_MAX;
            #pragma unroll
            for( int i = 0; i < S; ++i )
            {
                double x;
                if constexpr ( S == 2 ) { x = bt_seg_seg_d2( p[i], p[i], q[0], q[1] ); }
                else { const D3 c = bt_closest_on_triangle( p[i], q[0], q[1], q[2] ); const D3 d = d3sub( p[i], c ); x = d3dot( d, d ); }
                d2 = bt_min( d2, x );
            }
            #pragma unroll
            for( int i = 0; i < S; ++i )
            {
                double x;
                if constexpr ( S == 2 ) { x = bt_seg_seg_d2( q[i], q[i], p[0], p[1] ); }
                else { const D3 c = bt_closest_on_triangle( q[i], p[0], p[1], p[2] ); const D3 d = d3sub( q[i], c ); x = d3dot( d, d ); }
                d2 = bt_min( d2, x );
            }
            constexpr int E = ( S == 2 ) ? 1 : 3;
            #pragma unroll
            for( int i = 0; i < E; ++i )
            {
                #pragma unroll
                for( int j = 0; j < E; ++j ) { d2 = bt_min( d2, bt_seg_seg_d2( p[i], p[(i + 1) % S], q[j], q[(j + 1) % S] ) ); }
            }
            return d2;
        }

        struct ClassArgs
        {
            int64_t n_blocks;
            const uint64_t * blocks;      // sorted unique (I << 32 | J), I <= J
            const int32_t * leaf_b, * leaf_e;
            const double * tser;
            const int32_t * pre_cols;
            double near_theta2, inter_theta2;
            uint64_t * vf; int64_t vf_cap;
            uint64_t * inter; int64_t inter_cap;
            unsigned long long * ctr;     // [0] very near, [1] inter
        };

        // 0 near, 1 very near, 2 intersecting (BlockSplit_Kernel::ComputeLeaf as the host restates it)
        template<int S>
        __device__ int bt_classify( const ClassArgs & A, int32_t i, int32_t j )
        {
            _Pragma("clang fp contract(off)")
            constexpr int PS = 4 + 3 * S;
            const int32_t * a = A.pre_cols + S * (int64_t)i;
            const int32_t * b = A.pre_cols + S * (int64_t)j;
            #pragma unroll
            for( int k = 0; k < S; ++k )
            {
                #pragma unroll
                for( int l = 0; l < S; ++l ) { if( a[k] == b[l] ) { return 0; } }
            }
            const double * P = A.tser + PS * (int64_t)i;
            const double * Q = A.tser + PS * (int64_t)j;
            const double r2max = bt_max( P[0], Q[0] );
            const double r2min = bt_min( P[0], Q[0] );
            double c2 = 0.0;
            #pragma unroll
            for( int k = 0; k < 3; ++k ) { const double d = P[1 + k] - Q[1 + k]; c2 += d * d; }
            const double gap = sqrt( c2 ) - sqrt( P[0] ) - sqrt( Q[0] );
            if( gap > 0 && r2max < A.near_theta2 * gap * gap ) { return 0; }
            D3 p[S], q[S];
            #pragma unroll
            for( int k = 0; k < S; ++k )
            {
                p[k] = { P[4 + 3 * k], P[5 + 3 * k], P[6 + 3 * k] };
                q[k] = { Q[4 + 3 * k], Q[5 + 3 * k], Q[6 + 3 * k] };
            }
            const double d2 = bt_simplex_d2<S>( p, q );
            if( r2max < A.near_theta2 * d2 ) { return 0; }
            const double eps2 = DBL_EPSILON;
            return ( A.inter_theta2 * d2 <= eps2 * r2min ) ? 2 : 1;
        }

        // One wave per leaf block, one lane per primitive pair.
        template<int S>
        __global__ __launch_bounds__(BT_WG) void k_bt_classify( ClassArgs A )
        {
            const int64_t blk = (int64_t)blockIdx.x * ( BT_WG / 64 ) + ( threadIdx.x >> 6 );
            const int lane = threadIdx.x & 63;
            if( blk >= A.n_blocks ) { return; }
            const uint64_t key = A.blocks[blk];
            const int32_t I = (int32_t)( key >> 32 ), J = (int32_t)( key & 0xffffffffu );
            const int32_t ib = A.leaf_b[I], ni = A.leaf_e[I] - ib;
            const int32_t jb = A.leaf_b[J], nj = A.leaf_e[J] - jb;
            const int32_t np = ni * nj;
            for( int32_t base = 0; base < np; base += 64 )
            {
                const int32_t pidx = base + lane;
                int c = 0;
                int32_t a = 0, b = 0;
                if( pidx < np )
                {
                    const int32_t ii = pidx / nj, jj = pidx - ii * nj;
                    a = ib + ii; b = jb + jj;
                    if( I != J || ii < jj ) { c = bt_classify<S>( A, a, b ); }
                }
                const uint64_t pk = ( (uint64_t)(uint32_t)a << 32 ) | (uint64_t)(uint32_t)b;
                int e1, e2;
                const unsigned long long b1 = bt_wave_reserve( c == 1 ? 1 : 0, A.ctr + 0, lane, e1 );
                const unsigned long long b2 = bt_wave_reserve( c == 2 ? 1 : 0, A.ctr + 1, lane, e2 );
                if( c == 1 && b1 + e1 < (unsigned long long)A.vf_cap ) { A.vf[b1 + e1] = pk; }
                if( c == 2 && b2 + e2 < (unsigned long long)A.inter_cap ) { A.inter[b2 + e2] = pk; }
            }
        }

        // ---------------------------------------------------------------- tiles and CSR
        // Both orientations of the upper keys (I << 32 | J); the second copy of a diagonal key is UINT64_MAX.
        __global__ __launch_bounds__(BT_WG) void k_bt_sym_keys( int64_t n, const uint64_t * up, uint64_t * out, unsigned long long * ndiag )
        {
            const int64_t i = (int64_t)blockIdx.x * BT_WG + threadIdx.x;
            const int lane = threadIdx.x & 63;
            bool diag = false;
            if( i < n )
            {
                const uint64_t k = up[i];
                const uint64_t I = k >> 32, J = k & 0xffffffffu;
                diag = ( I == J );
                out[2 * i] = k;
                out[2 * i + 1] = diag ? ~0ull : ( ( J << 32 ) | I );
            }
            int ex;
            (void)bt_wave_reserve( diag ? 1 : 0, ndiag, lane, ex );
        }

        // rp[r] = first position of row r in the sorted keys, r = 0..rows.
        __global__ __launch_bounds__(BT_WG) void k_bt_row_ptr( int64_t rows, const uint64_t * keys, int64_t n, int64_t * rp )
        {
            const int64_t r = (int64_t)blockIdx.x * BT_WG + threadIdx.x;
            if( r > rows ) { return; }
            const uint64_t target = (uint64_t)r << 32;
            int64_t lo = 0, hi = n;
            while( lo < hi )
            {
                const int64_t m = lo + ( ( hi - lo ) >> 1 );
                if( keys[m] < target ) { lo = m + 1; } else { hi = m; }
            }
            rp[r] = lo;
        }

        __global__ __launch_bounds__(BT_WG) void k_bt_low_words( int64_t n, const uint64_t * keys, int32_t * ci )
        {
            const int64_t i = (int64_t)blockIdx.x * BT_WG + threadIdx.x;
            if( i < n ) { ci[i] = (int32_t)( keys[i] & 0xffffffffu ); }
        }

        __global__ __launch_bounds__(BT_WG) void k_bt_col_count( int64_t L, const int64_t * rp, const uint64_t * keys, const int32_t * leaf_b,
                                                                 const int32_t * leaf_e, int64_t * cnt )
        {
            const int64_t l = (int64_t)blockIdx.x * BT_WG + threadIdx.x;
            if( l > L ) { return; }
            if( l == L ) { cnt[L] = 0; return; }
            int64_t m = 0;
            for( int64_t k = rp[l]; k < rp[l + 1]; ++k ) { const int32_t J = (int32_t)( keys[k] & 0xffffffffu ); m += leaf_e[J] - leaf_b[J]; }
            cnt[l] = m;
        }

        __device__ __forceinline__ bool bt_contains( const uint64_t * a, int64_t n, uint64_t key )
        {
            int64_t lo = 0, hi = n;
            while( lo < hi )
            {
                const int64_t m = lo + ( ( hi - lo ) >> 1 );
                if( a[m] < key ) { lo = m + 1; } else { hi = m; }
            }
            return lo < n && a[lo] == key;
        }

        // One thread per (row leaf, neighbour leaf) entry: its columns and their row masks.
        __global__ __launch_bounds__(BT_WG) void k_bt_fill_tiles( int64_t n, const uint64_t * keys, const int64_t * rp, const int64_t * col_ptr,
                                                                  const int32_t * leaf_b, const int32_t * leaf_e,
                                                                  const uint64_t * vf, int64_t nvf, const uint64_t * inter, int64_t ninter,
                                                                  int32_t * col_prim, uint32_t * col_mask )
        {
            const int64_t e = (int64_t)blockIdx.x * BT_WG + threadIdx.x;
            if( e >= n ) { return; }
            const uint64_t k = keys[e];
            const int32_t l = (int32_t)( k >> 32 ), J = (int32_t)( k & 0xffffffffu );
            int64_t pos = col_ptr[l];
            for( int64_t f = rp[l]; f < e; ++f ) { const int32_t Jf = (int32_t)( keys[f] & 0xffffffffu ); pos += leaf_e[Jf] - leaf_b[Jf]; }
            const int32_t ib = leaf_b[l], ie = leaf_e[l];
            const bool any_excluded = ( nvf + ninter ) > 0;
            for( int32_t j = leaf_b[J]; j < leaf_e[J]; ++j )
            {
                uint32_t mask = 0;
                for( int32_t i = ib; i < ie; ++i )
                {
                    bool on = ( i != j );
                    if( on && any_excluded )
                    {
                        const uint64_t pk = ( (uint64_t)(uint32_t)min( i, j ) << 32 ) | (uint64_t)(uint32_t)max( i, j );
                        on = !( bt_contains( vf, nvf, pk ) || bt_contains( inter, ninter, pk ) );
                    }
                    if( on ) { mask |= ( 1u << ( i - ib ) ); }
                }
                col_prim[pos] = j;
                col_mask[pos] = mask;
                ++pos;
            }
        }
    } // namespace

    // ---------------------------------------------------------------- host driver
    namespace
    {
        // Sorts n keys from `in` into `out` (radix sort on all 64 bits) and, with unique, removes duplicates;
        // returns the number of keys in `out`.
        int64_t bt_sort_keys( BuiltTopology & B, uint64_t * in, uint64_t * out, int64_t n, bool unique, hipStream_t s )
        {
            if( n == 0 ) { return 0; }
            require( n < (int64_t)INT_MAX, "list too long for a 32-bit radix sort", RP_EUNSUPPORTED );
            size_t bytes = 0;
            RP_HIP( hipcub::DeviceRadixSort::SortKeys( nullptr, bytes, in, out, (int)n, 0, 64, s ) );
            B.tmp.require( bytes );
            RP_HIP( hipcub::DeviceRadixSort::SortKeys( B.tmp.get(), bytes, in, out, (int)n, 0, 64, s ) );
            if( !unique ) { return n; }
            size_t ubytes = 0;
            B.nsel.require( 1 );
            RP_HIP( hipcub::DeviceSelect::Unique( nullptr, ubytes, out, in, B.nsel.get(), (int)n, s ) );
            B.tmp.require( ubytes );
            RP_HIP( hipcub::DeviceSelect::Unique( B.tmp.get(), ubytes, out, in, B.nsel.get(), (int)n, s ) );
            int64_t m = 0;
            RP_HIP( hipMemcpyAsync( &m, B.nsel.get(), sizeof(int64_t), hipMemcpyDeviceToHost, s ) );
            RP_HIP( hipStreamSynchronize( s ) );
            RP_HIP( hipMemcpyAsync( out, in, (size_t)m * sizeof(uint64_t), hipMemcpyDeviceToDevice, s ) );
            return m;
        }

        template<int S>
        void build_tree( BuiltTopology & B, int64_t T, int thr, hipStream_t s )
        {
            const int64_t NC = 2 * T;   // node capacity (a binary tree over T leaves' worth of ranges)
            B.nb.require( NC ); B.ne.require( NC ); B.ndepth.require( NC ); B.nleft.require( NC ); B.nright.require( NC );
            B.nflag.require( NC ); B.ndir.require( NC ); B.nmid.require( NC ); B.nbox.require( 7 * NC ); B.nnl.require( NC ); B.nj0.require( NC );
            B.lvl_ind.require( T ); B.lvl_incl.require( T );
            B.ord.require( T ); B.ord2.require( T ); B.pnode.require( T ); B.pnode2.require( T );
            B.gcnt.require( T ); B.isl.require( T ); B.fb_list.require( T ); B.fb_ctr.require( 2 );
            const int32_t root[3] = { 0, (int32_t)T, 0 };
            RP_HIP( hipMemcpyAsync( B.nb.get(), &root[0], sizeof(int32_t), hipMemcpyHostToDevice, s ) );
            RP_HIP( hipMemcpyAsync( B.ne.get(), &root[1], sizeof(int32_t), hipMemcpyHostToDevice, s ) );
            RP_HIP( hipMemcpyAsync( B.ndepth.get(), &root[2], sizeof(int32_t), hipMemcpyHostToDevice, s ) );
            k_bt_iota<<<bt_grid( T ), BT_WG, 0, s>>>( T, B.ord.get(), B.pnode.get() );
            LevelArgs A{};
            A.nb = B.nb.get(); A.ne = B.ne.get(); A.ndepth = B.ndepth.get(); A.nleft = B.nleft.get(); A.nright = B.nright.get();
            A.nflag = B.nflag.get(); A.ndir = B.ndir.get(); A.nmid = B.nmid.get(); A.nbox = B.nbox.get(); A.nnl = B.nnl.get(); A.nj0 = B.nj0.get();
            A.lvl_ind = B.lvl_ind.get(); A.lvl_incl = B.lvl_incl.get(); A.gcnt = B.gcnt.get(); A.isl = B.isl.get();
            A.fb_list = B.fb_list.get(); A.fb_ctr = B.fb_ctr.get(); A.pser = B.pser.get(); A.thr = thr;
            int32_t * ord = B.ord.get(), * ord2 = B.ord2.get(), * pn = B.pnode.get(), * pn2 = B.pnode2.get();
            B.level_ptr.assign( 1, 0 );
            int64_t lvl0 = 0, n_lvl = 1;
            while( n_lvl > 0 )
            {
                require( lvl0 + n_lvl <= NC, "cluster tree node capacity exceeded" );
                A.lvl0 = lvl0; A.ord = ord; A.ord_new = ord2; A.pnode = pn; A.pnode_new = pn2;
                RP_HIP( hipMemsetAsync( B.fb_ctr.get(), 0, 2 * sizeof(unsigned int), s ) );
                k_bt_node_box<S><<<dim3( (unsigned)n_lvl ), BT_WG, 0, s>>>( A );
                k_bt_node_scan<S><<<dim3( (unsigned)n_lvl ), BT_WG, 0, s>>>( A );
                size_t scan_bytes = 0;
                RP_HIP( hipcub::DeviceScan::InclusiveSum( nullptr, scan_bytes, B.lvl_ind.get(), B.lvl_incl.get(), (int)n_lvl, s ) );
                B.tmp.require( scan_bytes );
                RP_HIP( hipcub::DeviceScan::InclusiveSum( B.tmp.get(), scan_bytes, B.lvl_ind.get(), B.lvl_incl.get(), (int)n_lvl, s ) );
                int32_t n_split = 0;
                unsigned int fb[2] = { 0, 0 };
                RP_HIP( hipMemcpyAsync( &n_split, B.lvl_incl.get() + ( n_lvl - 1 ), sizeof(int32_t), hipMemcpyDeviceToHost, s ) );
                RP_HIP( hipMemcpyAsync( fb, B.fb_ctr.get(), sizeof(fb), hipMemcpyDeviceToHost, s ) );
                RP_HIP( hipStreamSynchronize( s ) );
                require( fb[1] <= (unsigned)BT_SORT_MAX, "median split fallback on a cluster of more than 2048 primitives: build the tree on the host", RP_EUNSUPPORTED );
                const int64_t next0 = lvl0 + n_lvl;
                require( next0 + 2 * (int64_t)n_split <= NC, "cluster tree node capacity exceeded" );
                k_bt_children<<<bt_grid( n_lvl ), BT_WG, 0, s>>>( A, n_lvl, next0 );
                k_bt_permute<<<bt_grid( T ), BT_WG, 0, s>>>( A, T );
                if( fb[0] > 0 ) { k_bt_fallback_sort<S><<<dim3( fb[0] ), BT_WG, 0, s>>>( A ); }
                RP_HIP( hipGetLastError() );
                std::swap( ord, ord2 ); std::swap( pn, pn2 );
                lvl0 = next0;
                n_lvl = 2 * (int64_t)n_split;
                B.level_ptr.push_back( lvl0 );
            }
            // level_ptr: level d holds nodes [level_ptr[d], level_ptr[d+1])
            const int64_t N = lvl0;
            const int nlev = (int)B.level_ptr.size() - 1;
            B.C = N;
            B.depth = nlev - 1;
            B.final_ord = ord;
            B.desc.require( N ); B.dleaf.require( N ); B.g.require( N ); B.lb.require( N );
            for( int d = nlev - 1; d >= 0; --d )
            {
                const int64_t b = B.level_ptr[d], n = B.level_ptr[d + 1] - b;
                k_bt_desc<<<bt_grid( n ), BT_WG, 0, s>>>( b, n, B.nflag.get(), B.nleft.get(), B.nright.get(), B.desc.get(), B.dleaf.get() );
            }
            RP_HIP( hipMemsetAsync( B.g.get(), 0, sizeof(int32_t), s ) );
            RP_HIP( hipMemsetAsync( B.lb.get(), 0, sizeof(int32_t), s ) );
            for( int d = 0; d < nlev; ++d )
            {
                const int64_t b = B.level_ptr[d], n = B.level_ptr[d + 1] - b;
                k_bt_dfs<<<bt_grid( n ), BT_WG, 0, s>>>( b, n, B.nflag.get(), B.nleft.get(), B.nright.get(), B.desc.get(), B.dleaf.get(), B.g.get(), B.lb.get() );
            }
            int32_t L = 0;
            RP_HIP( hipMemcpyAsync( &L, B.dleaf.get(), sizeof(int32_t), hipMemcpyDeviceToHost, s ) );
            RP_HIP( hipStreamSynchronize( s ) );
            B.L = L;
            B.C_left.require( N ); B.C_right.require( N ); B.C_begin.require( N ); B.C_end.require( N ); B.C_depth.require( N );
            B.C_ser.require( 7 * N ); B.leaf_lookup.require( N );
            B.leaf_cluster.require( L ); B.leaf_b.require( L ); B.leaf_e.require( L );
            TreeOut O{ B.C_left.get(), B.C_right.get(), B.C_begin.get(), B.C_end.get(), B.C_depth.get(), B.C_ser.get(),
                       B.leaf_cluster.get(), B.leaf_lookup.get(), B.leaf_b.get(), B.leaf_e.get() };
            k_bt_scatter<<<bt_grid( N ), BT_WG, 0, s>>>( N, B.nb.get(), B.ne.get(), B.ndepth.get(), B.nflag.get(), B.nleft.get(), B.nright.get(),
                                                         B.nbox.get(), B.g.get(), B.lb.get(), O );
            RP_HIP( hipGetLastError() );
        }

        template<int S>
        void build_bct( BuiltTopology & B, int64_t T, double far_theta2, double near_theta2, double inter_theta2, hipStream_t s )
        {
            constexpr int PS = 4 + 3 * S;
            B.tser.require( PS * T ); B.pre_cols.require( S * T );
            k_bt_tree_order<S><<<bt_grid( T ), BT_WG, 0, s>>>( T, B.final_ord, B.pser.get(), B.ssimp.get(), B.tser.get(), B.pre_cols.get() );
            // traversal, one level per launch; the lists grow and the traversal restarts on overflow
            int64_t cap = std::max<int64_t>( (int64_t)1 << 20, 8 * B.C );
            int64_t far_cap = std::max<int64_t>( (int64_t)1 << 20, 16 * B.C );
            int64_t blk_cap = std::max<int64_t>( (int64_t)1 << 20, 32 * B.L );
            B.ctr.require( 4 );
            int64_t n_far = 0, n_blk = 0;
            for( int attempt = 0; ; ++attempt )
            {
                require( attempt < 8, "block cluster tree traversal lists exceed their limits", RP_ENOMEM );
                B.front_a.require( cap ); B.front_b.require( cap ); B.far_pairs.require( far_cap ); B.blk_pairs.require( blk_cap );
                int2 * in = B.front_a.get(), * out = B.front_b.get();
                const int2 root = make_int2( 0, 0 );
                RP_HIP( hipMemcpyAsync( in, &root, sizeof(int2), hipMemcpyHostToDevice, s ) );
                RP_HIP( hipMemsetAsync( B.ctr.get(), 0, 4 * sizeof(unsigned long long), s ) );
                int64_t n_in = 1;
                bool overflow = false;
                while( n_in > 0 )
                {
                    RP_HIP( hipMemsetAsync( B.ctr.get(), 0, sizeof(unsigned long long), s ) );
                    TravArgs A{ in, n_in, out, cap, B.far_pairs.get(), far_cap, B.blk_pairs.get(), blk_cap, B.ctr.get(),
                                B.C_left.get(), B.C_right.get(), B.leaf_lookup.get(), B.C_ser.get(), far_theta2 };
                    k_bt_traverse<<<bt_grid( n_in ), BT_WG, 0, s>>>( A );
                    RP_HIP( hipGetLastError() );
                    unsigned long long c[4];
                    RP_HIP( hipMemcpyAsync( c, B.ctr.get(), sizeof(c), hipMemcpyDeviceToHost, s ) );
                    RP_HIP( hipStreamSynchronize( s ) );
                    if( c[3] ) { overflow = true; break; }
                    n_in = (int64_t)c[0]; n_far = (int64_t)c[1]; n_blk = (int64_t)c[2];
                    std::swap( in, out );
                }
                if( !overflow ) { break; }
                cap *= 2; far_cap *= 2; blk_cap *= 2;
            }
            // sorted unique far pairs and leaf blocks
            B.keys_a.require( std::max<int64_t>( 2 * n_far, 2 * n_blk ) + 1 );
            B.keys_b.require( std::max<int64_t>( 2 * n_far, 2 * n_blk ) + 1 );
            B.far_up.require( n_far + 1 ); B.blocks.require( n_blk + 1 );
            if( n_far ) { k_bt_pair_keys<<<bt_grid( n_far ), BT_WG, 0, s>>>( n_far, B.far_pairs.get(), B.keys_a.get() ); }
            B.n_far_up = bt_sort_keys( B, B.keys_a.get(), B.far_up.get(), n_far, true, s );
            if( n_blk ) { k_bt_pair_keys<<<bt_grid( n_blk ), BT_WG, 0, s>>>( n_blk, B.blk_pairs.get(), B.keys_a.get() ); }
            B.n_blocks = bt_sort_keys( B, B.keys_a.get(), B.blocks.get(), n_blk, true, s );
            // classification of the primitive pairs inside the leaf blocks
            int64_t vf_cap = 1 << 16, inter_cap = 1 << 16;
            unsigned long long cc[2] = { 0, 0 };
            for( int attempt = 0; attempt < 2; ++attempt )
            {
                B.vf_raw.require( vf_cap ); B.inter_raw.require( inter_cap );
                RP_HIP( hipMemsetAsync( B.ctr.get(), 0, 2 * sizeof(unsigned long long), s ) );
                ClassArgs A{ B.n_blocks, B.blocks.get(), B.leaf_b.get(), B.leaf_e.get(), B.tser.get(), B.pre_cols.get(), near_theta2, inter_theta2,
                             B.vf_raw.get(), vf_cap, B.inter_raw.get(), inter_cap, B.ctr.get() };
                if( B.n_blocks ) { k_bt_classify<S><<<dim3( (unsigned)( ( B.n_blocks + 3 ) / 4 ) ), BT_WG, 0, s>>>( A ); }
                RP_HIP( hipGetLastError() );
                RP_HIP( hipMemcpyAsync( cc, B.ctr.get(), sizeof(cc), hipMemcpyDeviceToHost, s ) );
                RP_HIP( hipStreamSynchronize( s ) );
                if( (int64_t)cc[0] <= vf_cap && (int64_t)cc[1] <= inter_cap ) { break; }
                require( attempt == 0, "very-near list overflow", RP_ENOMEM );
                vf_cap = std::max<int64_t>( vf_cap, (int64_t)cc[0] ); inter_cap = std::max<int64_t>( inter_cap, (int64_t)cc[1] );
            }
            B.n_vf = (int64_t)cc[0]; B.n_inter = (int64_t)cc[1];
            B.vf.require( B.n_vf + 1 ); B.inter.require( B.n_inter + 1 );
            B.n_vf = bt_sort_keys( B, B.vf_raw.get(), B.vf.get(), B.n_vf, false, s );
            B.n_inter = bt_sort_keys( B, B.inter_raw.get(), B.inter.get(), B.n_inter, false, s );
            // near tiles: symmetric leaf-block graph, column counts, columns and masks
            const int64_t L = B.L;
            unsigned long long ndiag = 0;
            RP_HIP( hipMemsetAsync( B.ctr.get(), 0, sizeof(unsigned long long), s ) );
            if( B.n_blocks ) { k_bt_sym_keys<<<bt_grid( B.n_blocks ), BT_WG, 0, s>>>( B.n_blocks, B.blocks.get(), B.keys_a.get(), B.ctr.get() ); }
            RP_HIP( hipMemcpyAsync( &ndiag, B.ctr.get(), sizeof(ndiag), hipMemcpyDeviceToHost, s ) );
            RP_HIP( hipStreamSynchronize( s ) );
            B.nbr.require( 2 * B.n_blocks + 1 );
            bt_sort_keys( B, B.keys_a.get(), B.nbr.get(), 2 * B.n_blocks, false, s );
            B.n_nbr = 2 * B.n_blocks - (int64_t)ndiag;
            B.nbr_rp.require( L + 1 ); B.col_cnt.require( L + 1 ); B.col_ptr.require( L + 1 );
            k_bt_row_ptr<<<bt_grid( L + 1 ), BT_WG, 0, s>>>( L, B.nbr.get(), B.n_nbr, B.nbr_rp.get() );
            k_bt_col_count<<<bt_grid( L + 1 ), BT_WG, 0, s>>>( L, B.nbr_rp.get(), B.nbr.get(), B.leaf_b.get(), B.leaf_e.get(), B.col_cnt.get() );
            size_t sbytes = 0;
            RP_HIP( hipcub::DeviceScan::ExclusiveSum( nullptr, sbytes, B.col_cnt.get(), B.col_ptr.get(), (int)( L + 1 ), s ) );
            B.tmp.require( sbytes );
            RP_HIP( hipcub::DeviceScan::ExclusiveSum( B.tmp.get(), sbytes, B.col_cnt.get(), B.col_ptr.get(), (int)( L + 1 ), s ) );
            RP_HIP( hipMemcpyAsync( &B.n_cols, B.col_ptr.get() + L, sizeof(int64_t), hipMemcpyDeviceToHost, s ) );
            RP_HIP( hipStreamSynchronize( s ) );
            B.col_prim.require( B.n_cols + 1 ); B.col_mask.require( B.n_cols + 1 );
            if( B.n_nbr )
            {
                k_bt_fill_tiles<<<bt_grid( B.n_nbr ), BT_WG, 0, s>>>( B.n_nbr, B.nbr.get(), B.nbr_rp.get(), B.col_ptr.get(), B.leaf_b.get(), B.leaf_e.get(),
                                                                      B.vf.get(), B.n_vf, B.inter.get(), B.n_inter, B.col_prim.get(), B.col_mask.get() );
            }
            // far field: full symmetric CSR over clusters
            const int64_t nf = B.n_far_up;
            B.far_rp.require( B.C + 1 ); B.far_ci.require( 2 * nf + 1 );
            if( nf ) { k_bt_sym_keys<<<bt_grid( nf ), BT_WG, 0, s>>>( nf, B.far_up.get(), B.keys_a.get(), B.ctr.get() ); }
            bt_sort_keys( B, B.keys_a.get(), B.keys_b.get(), 2 * nf, false, s );
            k_bt_row_ptr<<<bt_grid( B.C + 1 ), BT_WG, 0, s>>>( B.C, B.keys_b.get(), 2 * nf, B.far_rp.get() );
            if( nf ) { k_bt_low_words<<<bt_grid( 2 * nf ), BT_WG, 0, s>>>( 2 * nf, B.keys_b.get(), B.far_ci.get() ); }
            B.far_nnz = 2 * nf;
            RP_HIP( hipGetLastError() );
            RP_HIP( hipStreamSynchronize( s ) );
        }
    } // namespace

    void build_topology( BuiltTopology & B, int dom_dim, int64_t V, int64_t T, const double * V_coords, const int32_t * simplices,
                         int split_threshold, double far_theta2, double near_theta2, double inter_theta2, hipStream_t s )
    {
        const int S = dom_dim + 1;
        B.coords.upload( V_coords, (size_t)( 3 * V ), s );
        B.simp.upload( simplices, (size_t)( S * T ), s );
        B.pser.require( (size_t)( ( 4 + 3 * S ) * T ) ); B.ssimp.require( (size_t)( S * T ) );
        B.T = T; B.dom_dim = dom_dim;
        if( S == 3 )
        {
            k_bt_serialize<3><<<bt_grid( T ), BT_WG, 0, s>>>( T, B.coords.get(), B.simp.get(), B.pser.get(), B.ssimp.get() );
            build_tree<3>( B, T, split_threshold, s );
            build_bct<3>( B, T, far_theta2, near_theta2, inter_theta2, s );
        }
        else
        {
            k_bt_serialize<2><<<bt_grid( T ), BT_WG, 0, s>>>( T, B.coords.get(), B.simp.get(), B.pser.get(), B.ssimp.get() );
            build_tree<2>( B, T, split_threshold, s );
            build_bct<2>( B, T, far_theta2, near_theta2, inter_theta2, s );
        }
    }

} // namespace rp
