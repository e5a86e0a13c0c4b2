// Plain device-to-device copy kernel of known size, used to calibrate rocprofv3's FETCH_SIZE /
// WRITE_SIZE counters against bytes actually moved (bench.py --calibrate-copy).

#include <hip/hip_runtime.h>
#include <cstdint>

__global__ __launch_bounds__(256) void k_calib_copy( const double4 * __restrict__ src, double4 * __restrict__ dst, int64_t n )
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for( int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride ) { dst[i] = src[i]; }
}

extern "C" int rpc_copy( int64_t bytes, int reps, double * ms_per_copy )
{
    const int64_t n = bytes / (int64_t)sizeof(double4);
    double4 * a = nullptr;
    double4 * b = nullptr;
    if( hipMalloc( (void**)&a, n * sizeof(double4) ) != hipSuccess ) { return 2; }
    if( hipMalloc( (void**)&b, n * sizeof(double4) ) != hipSuccess ) { (void)hipFree(a); return 2; }
    (void)hipMemset( a, 0, n * sizeof(double4) );
    hipEvent_t e0, e1;
    (void)hipEventCreate( &e0 ); (void)hipEventCreate( &e1 );
    const dim3 grid( 8192 ), block( 256 );
    hipLaunchKernelGGL( k_calib_copy, grid, block, 0, 0, a, b, n );
    (void)hipEventRecord( e0, 0 );
    for( int r = 0; r < reps; ++r ) { hipLaunchKernelGGL( k_calib_copy, grid, block, 0, 0, a, b, n ); }
    (void)hipEventRecord( e1, 0 );
    (void)hipEventSynchronize( e1 );
    float ms = 0;
    (void)hipEventElapsedTime( &ms, e0, e1 );
    if( ms_per_copy ) { *ms_per_copy = ms / reps; }
    (void)hipEventDestroy( e0 ); (void)hipEventDestroy( e1 );
    (void)hipFree( a ); (void)hipFree( b );
    return (hipGetLastError() == hipSuccess) ? 0 : 2;
}
