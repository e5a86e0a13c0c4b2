// Exact diagonal of the assembled tangent-point metric A (V x V, the same for the three columns) and
// the diagonal term of the two-level preconditioner P = P0 + w diag(A)^-1.
//
// A = Post (M_near + M_far) Pre with Pre = pre_vals (row 0: a/(n+1) per vertex, rows 1..3: a times
// the surface-gradient operator) and Post = Pre^T. For a vertex v with incident primitives i, j:
//   A_vv = sum_{i,j containing v} [ w_i0 w_j0 K0_ij + (w_i1 w_j1 + w_i2 w_j2 + w_i3 w_j3) K1_ij ]
// where K_ii = near diagonal of i + sum of the far diagonals of all clusters containing i, and
// for i != j (mesh-adjacent, hence always a near pair and never in an admissible far block, whose
// bounding boxes are disjoint) K_ij = near value (i, j) + sum of the far diagonals of all clusters
// containing both i and j (the root-to-LCA path).

#include "rp_launch.hpp"

namespace rp
{
    // S[2c + t] = sum of ff_diag[2a + t] over the clusters a on the path from c to the root.
    __global__ void k_path_diag( int64_t C, const int32_t * parent, const double * ff_diag, double * S )
    {
        const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if( c >= C ) { return; }
        double s0 = 0.0, s1 = 0.0;
        int depth = 0;
        for( int32_t a = (int32_t)c; a >= 0 && depth < 128; a = parent[a], ++depth ) { s0 += ff_diag[2 * a]; s1 += ff_diag[2 * a + 1]; }
        S[2 * c] = s0; S[2 * c + 1] = s1;
    }

    template<typename VT>
    __global__ __launch_bounds__(256) void k_metric_diag( DiagArgs A, const VT * nf_vals, double * dA )
    {
        const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if( v >= A.V ) { return; }
        const int size = A.size;
        double acc = 0.0;
        for( int64_t e1 = A.vp_ptr[v]; e1 < A.vp_ptr[v + 1]; ++e1 )
        {
            const int32_t i = A.vp_entry[e1] / size, s1 = A.vp_entry[e1] % size;
            const double * pi = A.pre_vals + (int64_t)4 * size * i + s1;
            const double a0 = pi[0], a1 = pi[size], a2 = pi[2 * size], a3 = pi[3 * size];
            const int32_t I = A.leaf_of_prim[i];
            const int32_t ib = A.leaf_begin[I];
            const int32_t nI = A.leaf_end[I] - ib;
            const int64_t cb = A.col_ptr[I];
            const int64_t ce = A.col_ptr[I + 1];
            for( int64_t e2 = A.vp_ptr[v]; e2 < A.vp_ptr[v + 1]; ++e2 )
            {
                const int32_t j = A.vp_entry[e2] / size, s2 = A.vp_entry[e2] % size;
                const double * pj = A.pre_vals + (int64_t)4 * size * j + s2;
                double k0 = 0.0, k1 = 0.0;
                if( i == j )
                {
                    k0 = A.nf_diag[2 * (int64_t)i]; k1 = A.nf_diag[2 * (int64_t)i + 1];
                    if( A.vf_diag ) { k0 += A.vf_diag[2 * (int64_t)i]; k1 += A.vf_diag[2 * (int64_t)i + 1]; }
                }
                else if( A.have_near )
                {
                    // column lists are sorted by primitive: binary search for j
                    int64_t lo = cb, hi = ce;
                    while( lo < hi )
                    {
                        const int64_t mid = (lo + hi) >> 1;
                        if( A.col_prim[mid] < j ) { lo = mid + 1; } else { hi = mid; }
                    }
                    if( lo < ce && A.col_prim[lo] == j && ((A.col_mask[lo] >> (i - ib)) & 1u) )
                    {
                        const int64_t slot = A.val_ptr[I] + (lo - cb) * nI + (i - ib);
                        k0 = (double)nf_vals[2 * slot]; k1 = (double)nf_vals[2 * slot + 1];
                    }
                }
                if( A.use_far )
                {
                    int32_t c = A.leaf_cluster[I];
                    int depth = 0;
                    while( !( A.C_begin[c] <= j && j < A.C_end[c] ) && depth < 128 ) { c = A.C_parent[c]; ++depth; }
                    k0 += A.S[2 * (int64_t)c]; k1 += A.S[2 * (int64_t)c + 1];
                }
                acc += a0 * pj[0] * k0 + ( a1 * pj[size] + a2 * pj[2 * size] + a3 * pj[3 * size] ) * k1;
            }
        }
        dA[v] = acc;
    }

    // Y += c * X / dA (per vertex, three columns)
    __global__ __launch_bounds__(256) void k_jacobi_add( int64_t V, double c, const double * X, int64_t ldX, const double * dA, double * Y, int64_t ldY )
    {
        const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if( v >= V ) { return; }
        const double s = dA[v] > 0.0 ? c / dA[v] : 0.0;
        for( int k = 0; k < 3; ++k ) { Y[ldY * v + k] += s * X[ldX * v + k]; }
    }

    void launch_metric_diag( bool mixed, const DiagArgs & A, const void * nf_vals, double * S, double * dA, hipStream_t s )
    {
        DiagArgs B = A;
        B.S = S;
        if( A.use_far && A.C > 0 )
        {
            hipLaunchKernelGGL( k_path_diag, dim3((unsigned)((A.C + 255) / 256)), dim3(256), 0, s, A.C, A.C_parent, A.ff_diag, S );
        }
        if( A.V <= 0 ) { return; }
        const dim3 grid( (unsigned)((A.V + 255) / 256) );
        if( mixed ) { hipLaunchKernelGGL( (k_metric_diag<float>), grid, dim3(256), 0, s, B, (const float *)nf_vals, dA ); }
        else        { hipLaunchKernelGGL( (k_metric_diag<double>), grid, dim3(256), 0, s, B, (const double *)nf_vals, dA ); }
    }

    void launch_jacobi_add( int64_t V, double c, const double * X, int64_t ldX, const double * dA, double * Y, int64_t ldY, hipStream_t s )
    {
        if( V > 0 ) { hipLaunchKernelGGL( k_jacobi_add, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, s, V, c, X, ldX, dA, Y, ldY ); }
    }

} // namespace rp
