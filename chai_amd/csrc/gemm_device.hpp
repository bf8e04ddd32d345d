// gemm_device.hpp -- shared device helpers of the GEMM family.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace chai_hip {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));
// Native 16-byte vectors for staging through registers and LDS (HIP's uint4 /
// float4 are structs with a union, which the compiler may keep in scratch).
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

// Round-to-nearest-even f32 -> bf16 bits (finite inputs; NaN kept quiet).
__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf16_bits_to_f32(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// a = hi + lo + r with hi = bf16(a), lo = bf16(a - hi); |r| <= 2^-16 |a|
// (8 significant bits each: |a - hi| <= 2^-8 |a|, |r| <= 2^-8 |a - hi|).
__device__ __forceinline__ void split_bf16(float a, uint16_t& hi, uint16_t& lo)
{
    hi = f32_to_bf16_rne(a);
    const float rem = a - bf16_bits_to_f32(hi);   // exact (Sterbenz)
    lo = f32_to_bf16_rne(rem);
}

typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));

// The same split for two values with gfx950's v_cvt_pk_bf16_f32 (round to
// nearest even, equal to f32_to_bf16_rne on finite inputs): packed hi and lo
// words, element 0 in the low half.
__device__ __forceinline__ void split_bf16x2(float a0, float a1, uint32_t& hi, uint32_t& lo)
{
    const f32x2_t v = {a0, a1};
    const bf16x2_t h = __builtin_convertvector(v, bf16x2_t);
    hi = __builtin_bit_cast(uint32_t, h);
    const f32x2_t hf = {__uint_as_float(hi << 16), __uint_as_float(hi & 0xffff0000u)};
    const f32x2_t r = v - hf;                     // exact (Sterbenz)
    lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(r, bf16x2_t));
}

typedef __attribute__((address_space(1))) void global_void_t;
typedef __attribute__((address_space(3))) void lds_void_t;

// Global-memory accesses through pointers the compiler cannot place (read
// from descriptors in memory): without the address space they become flat
// instructions, which also count against lgkmcnt and so make LDS waits wait
// for them.
template <typename T>
__device__ __forceinline__ T gload(const T* p)
{
    return *(const __attribute__((address_space(1))) T*)p;
}
template <typename T>
__device__ __forceinline__ void gstore(T* p, T v)
{
    *(__attribute__((address_space(1))) T*)p = v;
}

// 16 bytes per lane from global memory into LDS (global_load_lds_dwordx4):
// lane i of the wave writes bytes [16 i, 16 i + 16) after dst (dst is
// wave-uniform).  Completion is counted by vmcnt.
__device__ __forceinline__ void lds_dma16(const void* src, void* dst)
{
    __builtin_amdgcn_global_load_lds((global_void_t*)src, (lds_void_t*)dst, 16, 0, 0);
}

__device__ __forceinline__ void wait_lds_dma() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// XCD-aware tile order: the dispatcher hands consecutive workgroups to the 8
// XCDs round-robin; remap so that the tiles one XCD runs together are
// neighbours (shared A row panels / B column panels stay in that XCD's L2),
// then walk tiles in groups of GROUP_M tile rows.
template <int GROUP_M>
__device__ __forceinline__ void tile_coords(uint32_t bid, uint32_t tiles_m, uint32_t tiles_n,
                                            uint32_t& tm, uint32_t& tn)
{
    const uint32_t total = tiles_m * tiles_n;
    uint32_t id = bid;
    if ((total & 7u) == 0) {
        const uint32_t per_xcd = total >> 3;
        id = (bid & 7u) * per_xcd + (bid >> 3);
    }
    const uint32_t group_tiles = GROUP_M * tiles_n;
    const uint32_t group = id / group_tiles;
    const uint32_t first_m = group * GROUP_M;
    const uint32_t gm = (tiles_m - first_m) < GROUP_M ? (tiles_m - first_m) : GROUP_M;
    const uint32_t in_group = id - group * group_tiles;
    tm = first_m + in_group % gm;
    tn = in_group / gm;
}

}  // namespace chai_hip
