// Per-state data on the device from new vertex coordinates: the device side of
// SimplicialMesh::SemiStaticUpdate (reference src/SimplicialMesh.hpp:354-376).
//
//   k_near_far_rows   SimplexDataKernel ReadSimplex/WriteNear/WriteFar (src/SimplexDataKernel.hpp:92-238)
//                     for every tree position (ComputeNearFarData + ClusterTree::ComputePrimitiveData,
//                     src/SimplicialMesh/ComputeNearFarData.hpp:3-27, src/ClusterTree.hpp:319-338)
//   k_cluster_leaves  ComputeClusterData_Step on leaf clusters: area-weighted sums of the primitive far
//   k_cluster_level   rows, then children averaged into parents, one depth level per launch
//                     (src/ClusterTree/ComputeClusterData_Thread.hpp:50-129)
//
// The arithmetic follows the host restatement operation by operation with floating-point contraction
// off, so the rows equal the host's bit for bit; the existing preparation kernels then derive the SoA
// arrays exactly as they do from uploaded rows.

#include "rp_launch.hpp"

namespace rp
{
    namespace
    {
        inline dim3 upd_grid( int64_t n ) { return dim3( (unsigned)( ( n + 255 ) / 256 ) ); }
    }

    template<int N>
    __global__ __launch_bounds__(256) void k_near_far_rows( int64_t T, const int32_t * P_ordering, const int32_t * S,
                                                            const double * V, int64_t ldV, const double * charges,
                                                            double * near, double * far )
    {
        _Pragma("clang fp contract(off)")
        constexpr int SIZE = N + 1;
        constexpr int ND = 1 + 3 * SIZE + 6;
        const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
        if( i >= T ) { return; }
        const int64_t t = P_ordering[i];
        int32_t sx[SIZE], ss[SIZE];
        #pragma unroll
        for( int l = 0; l < SIZE; ++l ) { sx[l] = S[SIZE * t + l]; ss[l] = sx[l]; }
        // sort the vertex ids (at most three)
        #pragma unroll
        for( int a = 0; a < SIZE; ++a )
        {
            #pragma unroll
            for( int b = 0; b + 1 < SIZE - a; ++b )
            {
                if( ss[b + 1] < ss[b] ) { const int32_t tmp = ss[b]; ss[b] = ss[b + 1]; ss[b + 1] = tmp; }
            }
        }
        double hull[SIZE][3];
        #pragma unroll
        for( int l = 0; l < SIZE; ++l ) { for( int k = 0; k < 3; ++k ) { hull[l][k] = V[ldV * (int64_t)sx[l] + k]; } }
        double center[3];
        for( int k = 0; k < 3; ++k ) { center[k] = hull[0][k]; }
        double charge = charges ? charges[sx[0]] : 1.0;
        #pragma unroll
        for( int l = 1; l < SIZE; ++l )
        {
            for( int k = 0; k < 3; ++k ) { center[k] += hull[l][k]; }
            charge += charges ? charges[sx[l]] : 1.0;
        }
        const double nth = 1.0 / SIZE;
        for( int k = 0; k < 3; ++k ) { center[k] *= nth; }
        charge *= nth;
        double a = charge * ( N == 2 ? 0.5 : 1.0 );
        double df[3][N];
        #pragma unroll
        for( int l = 0; l < N; ++l )
        {
            for( int k = 0; k < 3; ++k ) { df[k][l] = V[ldV * (int64_t)ss[l + 1] + k] - V[ldV * (int64_t)ss[0] + k]; }
        }
        double g[N][N];
        #pragma unroll
        for( int r = 0; r < N; ++r )
        {
            #pragma unroll
            for( int c = 0; c < N; ++c )
            {
                double s = 0;
                for( int k = 0; k < 3; ++k ) { s += df[k][r] * df[k][c]; }
                g[r][c] = s;
            }
        }
        double U[N][N];
        #pragma unroll
        for( int r = 0; r < N; ++r ) { for( int c = 0; c < N; ++c ) { U[r][c] = 0.0; } }
        #pragma unroll
        for( int r = 0; r < N; ++r )
        {
            double d = g[r][r];
            for( int k = 0; k < r; ++k ) { d -= U[k][r] * U[k][r]; }
            U[r][r] = sqrt( d );
            #pragma unroll
            for( int c = r + 1; c < N; ++c )
            {
                double s = g[r][c];
                for( int k = 0; k < r; ++k ) { s -= U[k][r] * U[k][c]; }
                U[r][c] = s / U[r][r];
            }
        }
        #pragma unroll
        for( int l = 0; l < N; ++l ) { a *= U[l][l]; }
        double dd[N][3];
        #pragma unroll
        for( int k = 0; k < 3; ++k )
        {
            double y[N];
            #pragma unroll
            for( int r = 0; r < N; ++r )
            {
                double s = df[k][r];
                for( int q = 0; q < r; ++q ) { s -= U[q][r] * y[q]; }
                y[r] = s / U[r][r];
            }
            #pragma unroll
            for( int r = N - 1; r >= 0; --r )
            {
                double s = y[r];
                for( int q = r + 1; q < N; ++q ) { s -= U[r][q] * dd[q][k]; }
                dd[r][k] = s / U[r][r];
            }
        }
        double P[3][3];
        #pragma unroll
        for( int r = 0; r < 3; ++r )
        {
            #pragma unroll
            for( int c = 0; c < 3; ++c )
            {
                double s = 0;
                for( int l = 0; l < N; ++l ) { s += df[r][l] * dd[l][c]; }
                P[r][c] = ( r == c ? 1.0 : 0.0 ) - s;
            }
        }
        const int pi[6] = { 0, 0, 0, 1, 1, 2 }, pj[6] = { 0, 1, 2, 1, 2, 2 };
        double * nr = near + (int64_t)ND * i;
        nr[0] = a;
        #pragma unroll
        for( int l = 0; l < SIZE; ++l ) { for( int k = 0; k < 3; ++k ) { nr[1 + 3 * l + k] = hull[l][k]; } }
        #pragma unroll
        for( int k = 0; k < 6; ++k ) { nr[1 + 3 * SIZE + k] = P[pi[k]][pj[k]]; }
        double * fr = far + 10 * i;
        fr[0] = a;
        for( int k = 0; k < 3; ++k ) { fr[1 + k] = center[k]; }
        #pragma unroll
        for( int k = 0; k < 6; ++k ) { fr[4 + k] = P[pi[k]][pj[k]]; }
    }

    __global__ __launch_bounds__(256) void k_cluster_leaves( int64_t L, const int32_t * leaf_cluster, const int32_t * C_begin,
                                                             const int32_t * C_end, const double * pfar, int singletons, double * C_far )
    {
        _Pragma("clang fp contract(off)")
        const int64_t l = (int64_t)blockIdx.x * 256 + threadIdx.x;
        if( l >= L ) { return; }
        const int32_t c = leaf_cluster[l];
        const int32_t b = C_begin[c], e = C_end[c];
        double * C = C_far + 10 * (int64_t)c;
        if( singletons )
        {
            for( int j = 0; j < 10; ++j ) { C[j] = pfar[10 * (int64_t)b + j]; }
            return;
        }
        double acc[10];
        {
            const double * P = pfar + 10 * (int64_t)b;
            const double a = P[0];
            acc[0] = a;
            for( int j = 1; j < 10; ++j ) { acc[j] = a * P[j]; }
        }
        for( int32_t i = b + 1; i < e; ++i )
        {
            const double * P = pfar + 10 * (int64_t)i;
            const double a = P[0];
            acc[0] += a;
            for( int j = 1; j < 10; ++j ) { acc[j] += a * P[j]; }
        }
        const double inv = 1.0 / acc[0];
        C[0] = acc[0];
        for( int j = 1; j < 10; ++j ) { C[j] = acc[j] * inv; }
    }

    __global__ __launch_bounds__(256) void k_cluster_level( int64_t n, const int32_t * nodes, const int32_t * C_left,
                                                            const int32_t * C_right, double * C_far )
    {
        _Pragma("clang fp contract(off)")
        const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
        if( k >= n ) { return; }
        const int32_t c = nodes[k];
        const int32_t L = C_left[c];
        if( L < 0 ) { return; }
        const int32_t R = C_right[c];
        const double * CL = C_far + 10 * (int64_t)L;
        const double * CR = C_far + 10 * (int64_t)R;
        double * CC = C_far + 10 * (int64_t)c;
        double lw = CL[0], rw = CR[0];
        const double mass = lw + rw;
        CC[0] = mass;
        const double inv = 1.0 / mass;
        lw *= inv; rw *= inv;
        for( int j = 1; j < 10; ++j ) { CC[j] = lw * CL[j] + rw * CR[j]; }
    }

    void launch_near_far_rows( int n, int64_t T, const int32_t * P_ordering, const int32_t * S, const double * V, int64_t ldV,
                               const double * charges, double * near, double * far, hipStream_t s )
    {
        if( T <= 0 ) { return; }
        if( n == 2 ) { hipLaunchKernelGGL( k_near_far_rows<2>, upd_grid( T ), dim3(256), 0, s, T, P_ordering, S, V, ldV, charges, near, far ); }
        else         { hipLaunchKernelGGL( k_near_far_rows<1>, upd_grid( T ), dim3(256), 0, s, T, P_ordering, S, V, ldV, charges, near, far ); }
    }

    void launch_cluster_leaves( int64_t L, const int32_t * leaf_cluster, const int32_t * C_begin, const int32_t * C_end,
                                const double * pfar, int singletons, double * C_far, hipStream_t s )
    {
        if( L > 0 ) { hipLaunchKernelGGL( k_cluster_leaves, upd_grid( L ), dim3(256), 0, s, L, leaf_cluster, C_begin, C_end, pfar, singletons, C_far ); }
    }

    void launch_cluster_level( int64_t n, const int32_t * nodes, const int32_t * C_left, const int32_t * C_right, double * C_far, hipStream_t s )
    {
        if( n > 0 ) { hipLaunchKernelGGL( k_cluster_level, upd_grid( n ), dim3(256), 0, s, n, nodes, C_left, C_right, C_far ); }
    }
}
