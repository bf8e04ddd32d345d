// reduce_device.hpp -- wavefront (64-lane) and workgroup reductions.
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

namespace chai_hip {

// Butterfly over the 64 lanes of a wavefront; every lane ends with the total.
// The combination order is fixed, so results are deterministic.
__device__ __forceinline__ double wave_sum_f64(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ long long wave_sum_i64(long long v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ float wave_sum_f32(float v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Workgroup sum; the result is valid in thread 0 (and broadcast to all).
template <int THREADS>
__device__ __forceinline__ double block_sum_f64(double v)
{
    constexpr int WAVES = THREADS / 64;
    __shared__ double red[WAVES];
    v = wave_sum_f64(v);
    if (WAVES == 1) return v;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < WAVES; i++) t += red[i];
    __syncthreads();
    return t;
}

template <int THREADS>
__device__ __forceinline__ long long block_sum_i64(long long v)
{
    constexpr int WAVES = THREADS / 64;
    __shared__ long long redi[WAVES];
    v = wave_sum_i64(v);
    if (WAVES == 1) return v;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) redi[wave] = v;
    __syncthreads();
    long long t = 0;
#pragma unroll
    for (int i = 0; i < WAVES; i++) t += redi[i];
    __syncthreads();
    return t;
}

}  // namespace chai_hip
