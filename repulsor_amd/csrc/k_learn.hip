// Learned preconditioner P = D^{1/2} (P0 + w diag(A)^-1) D^{1/2}: a 2-layer MLP (width 64) per
// cluster-tree node maps node features to a positive scaling s_c = softplus(o_c); log d_v starts as the
// mean over the primitives i containing v of sum_{c on the path leaf(i) -> root} log s_c, which is
// constant per leaf, and is smoothed by LP_SMOOTH_STEPS steps of S = (vertex mean of primitive means) so
// that d carries no jumps at leaf boundaries; d_v = exp(log d_v).
// Inference runs with BF16 weights and features (FP32 accumulation); training computes the loss
// gradient in FP32/FP64 with fixed-order reductions (no atomics).

#include "rp_launch.hpp"

#include <hip/hip_bf16.h>

namespace rp
{
    __device__ __forceinline__ float softplus_f( float x ) { return x > 20.f ? x : fmaxf( log1pf( expf( x ) ), 1e-30f ); }
    __device__ __forceinline__ float sigmoid_f( float x ) { return 1.f / ( 1.f + expf( -x ) ); }

    // lam[c] = log softplus( w2 . tanh( W1 f_c + b1 ) + b2 ): the hidden layer (W1, b1 and the features)
    // in BF16, the output layer (w2, b2) in FP32. b2 carries the scalings' common offset (softplus^-1(1)
    // for the identity); rounded to BF16 it would shift every node's log s by about 1e-3, which sums to
    // a depth-dependent scaling of a few percent along the root paths.
    __global__ __launch_bounds__(256) void k_mlp_infer_bf16( int64_t C, int F, const float * feats, const __hip_bfloat16 * p, const float * p32, double * lam )
    {
        const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if( c >= C ) { return; }
        const __hip_bfloat16 * W1 = p;                     // MLP_H x F
        const __hip_bfloat16 * b1 = W1 + MLP_H * F;
        const float * w2 = p32 + MLP_H * F + MLP_H;
        const float * b2 = w2 + MLP_H;
        float f[MLP_FMAX];
        for( int k = 0; k < F; ++k ) { f[k] = __bfloat162float( __float2bfloat16( feats[(int64_t)F * c + k] ) ); }
        float o = b2[0];
        for( int j = 0; j < MLP_H; ++j )
        {
            float z = __bfloat162float( b1[j] );
            for( int k = 0; k < F; ++k ) { z += __bfloat162float( W1[j * F + k] ) * f[k]; }
            o += w2[j] * tanhf( z );
        }
        lam[c] = (double)logf( softplus_f( o ) );
    }

    // FP32 forward used by training (same formula, FP32 weights).
    __device__ __forceinline__ float mlp_forward_f32( int F, const float * f, const float * p, float * h )
    {
        const float * W1 = p; const float * b1 = W1 + MLP_H * F; const float * w2 = b1 + MLP_H; const float * b2 = w2 + MLP_H;
        float o = b2[0];
        for( int j = 0; j < MLP_H; ++j )
        {
            float z = b1[j];
            for( int k = 0; k < F; ++k ) { z += W1[j * F + k] * f[k]; }
            h[j] = tanhf( z );
            o += w2[j] * h[j];
        }
        return o;
    }

    __global__ __launch_bounds__(256) void k_mlp_infer_f32( int64_t C, int F, const float * feats, const float * p, double * lam )
    {
        const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if( c >= C ) { return; }
        float h[MLP_H];
        const float o = mlp_forward_f32( F, feats + (int64_t)F * c, p, h );
        lam[c] = (double)logf( softplus_f( o ) );
    }

    // sqrt_d[v] = exp( 0.5 * mean over entries of v of leaf_lam[leaf_of_prim[prim]] ), with leaf_lam the
    // root-path sums of lam.
    // logd[v] = mean over the primitives containing v of the leaf-to-root sum of their leaf.
    __global__ void k_vertex_logd( int64_t V, int size, const int64_t * vp_ptr, const int32_t * vp_entry, const int32_t * leaf_of_prim,
                                   const double * leaf_lam, double * logd )
    {
        const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if( v >= V ) { return; }
        double s = 0.0;
        const int64_t b = vp_ptr[v], e = vp_ptr[v + 1];
        for( int64_t k = b; k < e; ++k ) { s += leaf_lam[leaf_of_prim[vp_entry[k] / size]]; }
        logd[v] = ( e > b ) ? s / (double)( e - b ) : 0.0;
    }

    // Smoothing step S = (vertex mean of primitive values) o (primitive mean of vertex values), first half:
    // pm[i] = mean over the vertices of primitive i of x.
    __global__ void k_prim_mean( int64_t T, int size, const int32_t * pre_cols, const double * x, double * pm )
    {
        const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if( i >= T ) { return; }
        double s = 0.0;
        for( int l = 0; l < size; ++l ) { s += x[pre_cols[(int64_t)size * i + l]]; }
        pm[i] = s / (double)size;
    }

    // Second half of S: x[v] = mean over the primitives containing v of pm.
    __global__ void k_vertex_mean( int64_t V, int size, const int64_t * vp_ptr, const int32_t * vp_entry, const double * pm, double * x )
    {
        const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if( v >= V ) { return; }
        double s = 0.0;
        const int64_t b = vp_ptr[v], e = vp_ptr[v + 1];
        for( int64_t k = b; k < e; ++k ) { s += pm[vp_entry[k] / size]; }
        x[v] = ( e > b ) ? s / (double)( e - b ) : 0.0;
    }

    // Transpose of k_prim_mean: g[v] = (1 / size) * sum over the primitives containing v of h.
    __global__ void k_vertex_sum_prim( int64_t V, int size, const int64_t * vp_ptr, const int32_t * vp_entry, const double * h, double * g )
    {
        const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if( v >= V ) { return; }
        double s = 0.0;
        for( int64_t k = vp_ptr[v]; k < vp_ptr[v + 1]; ++k ) { s += h[vp_entry[k] / size]; }
        g[v] = s / (double)size;
    }

    __global__ void k_exp_half( int64_t V, const double * x, double * sqrt_d )
    {
        const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if( v >= V ) { return; }
        sqrt_d[v] = exp( 0.5 * x[v] );
    }

    // Y = s * X (row scaling by sqrt_d), optionally Y = s * X + beta * Y.
    __global__ __launch_bounds__(256) void k_row_scale( int64_t V, const double * s, const double * X, int64_t ldX, double alpha, double beta, double * Y, int64_t ldY )
    {
        const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if( v >= V ) { return; }
        const double c = alpha * s[v];
        for( int k = 0; k < 3; ++k )
        {
            const double val = c * X[ldX * v + k];
            Y[ldY * v + k] = ( beta == 0.0 ) ? val : val + beta * Y[ldY * v + k];
        }
    }

    // ---------------------------------------------------------------------------- training
    // Per vertex: g_logd[v] = 0.5 sqrt_d[v] ( <g_v, q_v> + <t_v, r_v> ), the derivative of the loss with
    // respect to log d_v (see rp_solve_capi.hip, learned_grad).
    __global__ void k_grad_logd( int64_t V, const double * sqrt_d, const double * g, const double * q, const double * t, const double * r, double * out )
    {
        const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if( v >= V ) { return; }
        double s = 0.0;
        for( int k = 0; k < 3; ++k ) { s += g[3 * v + k] * q[3 * v + k] + t[3 * v + k] * r[3 * v + k]; }
        out[v] = 0.5 * sqrt_d[v] * s;
    }

    // Per primitive (tree order): g_prim[i] = sum over its vertices v of g_logd[v] / (entries of v).
    __global__ void k_grad_prim( int64_t T, int size, const int32_t * pre_cols, const int64_t * vp_ptr, const double * g_logd, double * g_prim )
    {
        const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        if( i >= T ) { return; }
        double s = 0.0;
        for( int l = 0; l < size; ++l )
        {
            const int32_t v = pre_cols[(int64_t)size * i + l];
            s += g_logd[v] / (double)( vp_ptr[v + 1] - vp_ptr[v] );
        }
        g_prim[i] = s;
    }

    // Parameter gradient from dL/dlam per cluster: one partial per block, summed by k_param_reduce.
    __global__ __launch_bounds__(256) void k_mlp_backward( int64_t C, int F, const float * feats, const float * p, const double * g_lam, double * part )
    {
        __shared__ double red[256];
        const int NP = MLP_H * F + 2 * MLP_H + 1;
        const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        double go = 0.0;
        float h[MLP_H];
        float f[MLP_FMAX];
        if( c < C )
        {
            for( int k = 0; k < F; ++k ) { f[k] = feats[(int64_t)F * c + k]; }
            const float o = mlp_forward_f32( F, f, p, h );
            // lam = log softplus(o): dlam/do = sigmoid(o) / softplus(o)
            go = g_lam[c] * (double)( sigmoid_f( o ) / softplus_f( o ) );
        }
        const float * w2 = p + MLP_H * F + MLP_H;
        // each parameter: block sum of its per-cluster contribution
        for( int q = 0; q < NP; ++q )
        {
            double val = 0.0;
            if( c < C )
            {
                if( q < MLP_H * F )                 { const int j = q / F, k = q % F; val = go * w2[j] * (1.0 - (double)h[j] * h[j]) * f[k]; }
                else if( q < MLP_H * F + MLP_H )    { const int j = q - MLP_H * F; val = go * w2[j] * (1.0 - (double)h[j] * h[j]); }
                else if( q < MLP_H * F + 2 * MLP_H ){ const int j = q - MLP_H * F - MLP_H; val = go * h[j]; }
                else                                { val = go; }
            }
            red[threadIdx.x] = val;
            __syncthreads();
            for( int s2 = 128; s2 > 0; s2 >>= 1 )
            {
                if( (int)threadIdx.x < s2 ) { red[threadIdx.x] += red[threadIdx.x + s2]; }
                __syncthreads();
            }
            if( threadIdx.x == 0 ) { part[(int64_t)blockIdx.x * NP + q] = red[0]; }
            __syncthreads();
        }
    }

    // grad[q] = sum over blocks of part[b][q] (fixed order)
    __global__ void k_param_reduce( int64_t nb, int NP, const double * part, double * grad )
    {
        const int q = blockIdx.x * blockDim.x + threadIdx.x;
        if( q >= NP ) { return; }
        double s = 0.0;
        for( int64_t b = 0; b < nb; ++b ) { s += part[b * NP + q]; }
        grad[q] = s;
    }

    // Adam on FP32 parameters with FP64 gradients.
    __global__ void k_adam( int NP, float * p, float * m, float * v, const double * grad, float lr, float b1, float b2, float eps, int t )
    {
        const int q = blockIdx.x * blockDim.x + threadIdx.x;
        if( q >= NP ) { return; }
        const float g = (float)grad[q];
        m[q] = b1 * m[q] + (1.f - b1) * g;
        v[q] = b2 * v[q] + (1.f - b2) * g * g;
        const float mh = m[q] / (1.f - powf( b1, (float)t ));
        const float vh = v[q] / (1.f - powf( b2, (float)t ));
        p[q] -= lr * mh / ( sqrtf( vh ) + eps );
    }

    __global__ void k_to_bf16( int n, const float * src, __hip_bfloat16 * dst )
    {
        const int q = blockIdx.x * blockDim.x + threadIdx.x;
        if( q < n ) { dst[q] = __float2bfloat16( src[q] ); }
    }

    static dim3 g256( int64_t n ) { return dim3( (unsigned)std::max<int64_t>( 1, (n + 255) / 256 ) ); }

    void launch_mlp_infer( bool bf16, int64_t C, int F, const float * feats, const float * p32, void * p16, double * lam, hipStream_t s )
    {
        if( C <= 0 ) { return; }
        if( bf16 )
        {
            const int NP = MLP_H * F + 2 * MLP_H + 1;
            hipLaunchKernelGGL( k_to_bf16, g256( NP ), dim3(256), 0, s, NP, p32, (__hip_bfloat16 *)p16 );
            hipLaunchKernelGGL( k_mlp_infer_bf16, g256( C ), dim3(256), 0, s, C, F, feats, (const __hip_bfloat16 *)p16, p32, lam );
        }
        else { hipLaunchKernelGGL( k_mlp_infer_f32, g256( C ), dim3(256), 0, s, C, F, feats, p32, lam ); }
    }

    void launch_vertex_scale( int64_t V, int64_t T, int size, const int32_t * pre_cols, const int64_t * vp_ptr, const int32_t * vp_entry,
                              const int32_t * leaf_of_prim, const double * leaf_lam, double * logd, double * pm, double * sqrt_d, hipStream_t s )
    {
        if( V <= 0 ) { return; }
        hipLaunchKernelGGL( k_vertex_logd, g256( V ), dim3(256), 0, s, V, size, vp_ptr, vp_entry, leaf_of_prim, leaf_lam, logd );
        for( int k = 0; k < LP_SMOOTH_STEPS; ++k )
        {
            hipLaunchKernelGGL( k_prim_mean, g256( T ), dim3(256), 0, s, T, size, pre_cols, logd, pm );
            hipLaunchKernelGGL( k_vertex_mean, g256( V ), dim3(256), 0, s, V, size, vp_ptr, vp_entry, pm, logd );
        }
        hipLaunchKernelGGL( k_exp_half, g256( V ), dim3(256), 0, s, V, logd, sqrt_d );
    }

    void launch_smooth_transpose( int64_t V, int64_t T, int size, const int32_t * pre_cols, const int64_t * vp_ptr, const int32_t * vp_entry,
                                  double * g, double * h, hipStream_t s )
    {
        if( V <= 0 ) { return; }
        for( int k = 0; k < LP_SMOOTH_STEPS; ++k )
        {
            launch_grad_prim( T, size, pre_cols, vp_ptr, g, h, s );                                           // transpose of k_vertex_mean
            hipLaunchKernelGGL( k_vertex_sum_prim, g256( V ), dim3(256), 0, s, V, size, vp_ptr, vp_entry, h, g );   // transpose of k_prim_mean
        }
    }

    void launch_row_scale( int64_t V, const double * sd, const double * X, int64_t ldX, double alpha, double beta, double * Y, int64_t ldY, hipStream_t s )
    {
        if( V > 0 ) { hipLaunchKernelGGL( k_row_scale, g256( V ), dim3(256), 0, s, V, sd, X, ldX, alpha, beta, Y, ldY ); }
    }

    void launch_grad_logd( int64_t V, const double * sd, const double * g, const double * q, const double * t, const double * r, double * out, hipStream_t s )
    {
        if( V > 0 ) { hipLaunchKernelGGL( k_grad_logd, g256( V ), dim3(256), 0, s, V, sd, g, q, t, r, out ); }
    }

    void launch_grad_prim( int64_t T, int size, const int32_t * pre_cols, const int64_t * vp_ptr, const double * g_logd, double * g_prim, hipStream_t s )
    {
        if( T > 0 ) { hipLaunchKernelGGL( k_grad_prim, g256( T ), dim3(256), 0, s, T, size, pre_cols, vp_ptr, g_logd, g_prim ); }
    }

    int64_t mlp_backward_blocks( int64_t C ) { return std::max<int64_t>( 1, (C + 255) / 256 ); }

    void launch_mlp_backward( int64_t C, int F, const float * feats, const float * p, const double * g_lam, double * part, double * grad, hipStream_t s )
    {
        const int NP = MLP_H * F + 2 * MLP_H + 1;
        const int64_t nb = mlp_backward_blocks( C );
        hipLaunchKernelGGL( k_mlp_backward, dim3((unsigned)nb), dim3(256), 0, s, C, F, feats, p, g_lam, part );
        hipLaunchKernelGGL( k_param_reduce, g256( NP ), dim3(256), 0, s, nb, NP, part, grad );
    }

    void launch_adam( int NP, float * p, float * m, float * v, const double * grad, float lr, float b1, float b2, float eps, int t, hipStream_t s )
    {
        hipLaunchKernelGGL( k_adam, g256( NP ), dim3(256), 0, s, NP, p, m, v, grad, lr, b1, b2, eps, t );
    }

} // namespace rp
