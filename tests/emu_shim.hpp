// CPU emulation shim for emitted trace kernels (debugging only)
#pragma once
#include <cstdint>
#include <cstring>
#include <cmath>
#define __HIPCC_RTC__ 1
#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __shared__ static
#define __launch_bounds__(...)
struct EmuDim { unsigned x, y, z; };
extern thread_local EmuDim blockIdx, threadIdx, gridDim, blockDim;
inline void __syncthreads() {}
template <class T> inline T __shfl_xor(T v, int, int = 64) { return v; }
inline uint32_t __float_as_uint(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
inline float __uint_as_float(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
inline int64_t __double_as_longlong(double d) { int64_t u; memcpy(&u, &d, 8); return u; }
inline double __longlong_as_double(int64_t u) { double d; memcpy(&d, &u, 8); return d; }
inline uint32_t __umulhi(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) >> 32); }
inline uint64_t __umul64hi(uint64_t a, uint64_t b) { return (uint64_t)(((__uint128_t)a * b) >> 64); }
inline int __mulhi(int a, int b) { return (int)(((int64_t)a * b) >> 32); }
enum { CHAI_U32 = 0, CHAI_I32 = 1, CHAI_F32 = 4, CHAI_F64 = 8, CHAI_RNG_PHILOX = 1230, CHAI_RNG_THREEFRY = 1231 };
inline int __clz(int x) { return x ? __builtin_clz((unsigned)x) : 32; }
inline int __popc(unsigned x) { return __builtin_popcount(x); }
// v_bitop3_b32: bit i of the result is bit (a_i b_i c_i) of the truth table imm
inline uint32_t emu_bitop3(uint32_t a, uint32_t b, uint32_t c, unsigned imm)
{
    uint32_t r = 0;
    for (unsigned t = 0; t < 8; t++)
        if ((imm >> t) & 1u) r |= ((t & 4) ? a : ~a) & ((t & 2) ? b : ~b) & ((t & 1) ? c : ~c);
    return r;
}
#define __builtin_amdgcn_bitop3_b32(a, b, c, imm) emu_bitop3((a), (b), (c), (imm))
// v_rcp_f32 / v_sqrt_f32 are within an ulp on the device; the correctly rounded
// host operations are within that bound
inline float emu_rcpf(float x) { return 1.0f / x; }
#define __builtin_amdgcn_rcpf(x) emu_rcpf(x)
#define __builtin_amdgcn_sqrtf(x) sqrtf(x)
