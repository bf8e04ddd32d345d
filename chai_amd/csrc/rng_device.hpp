// rng_device.hpp -- Chai's counter-based RNG stream as gfx950 device code.
//
// Stream definition (shared with the CPU oracle, checked bit-exact on GPU):
//   - block functions Philox2x32/4x32/2x64-10 and Threefry2x32/4x32/2x64-20
//     (Random123 1.06 as Chai emits it: Random123/R123Philox.cpp:34-286,
//     Random123/R123Threefry.cpp:131-290; Chai's Threefry4x32 mixes word pairs
//     (0,1),(2,3) in every round, R123Threefry.cpp:238-255);
//   - key = {vidx, 0xdecafbad[, 0xfacebead, 0x12345678]},
//     ctr = {seed, 0xf00dcafe, 0xdeadbeef, 0xbeeff00d} (R123Philox.cpp:373-381,
//     R123Threefry.cpp:377-387); 64-bit: key = {vidx[, 0xdecafbad]}, ctr = {seed, 0xf00dcafe};
//   - float = (float)(int32)w * 2^-32 + 0.5; double = (double)(int64)w * 2^-64 + 0.5
//     (R123Philox.cpp:405-428);
//   - Box-Muller r = sqrt(-2 ln u0), [r sin(2 pi u1), r cos(2 pi u1), ...]
//     (R123Philox.cpp:480-507) evaluated in double with this file's own log and
//     sin/cos(2 pi u) series (no contraction, explicit fma), rounded once.
// All code here is compiled with -ffp-contract=off.
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <cstdint>
#endif

namespace chai_hip {

// ------------------------------------------------------------ block functions
// a ^ b ^ c in one v_bitop3_b32 (truth table 0x96); the two-XOR form costs two
// full-rate issue slots, this one about one and a quarter
__device__ __forceinline__ uint32_t xor3_u32(uint32_t a, uint32_t b, uint32_t c)
{
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        // one 32 x 32 -> 64-bit product per multiplier (v_mad_u64_u32, half rate)
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
        const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
        const uint32_t n0 = xor3_u32(hi1, c[1], k0), n2 = xor3_u32(hi0, c[3], k1);
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    }
}

__device__ __forceinline__ void philox2x32_10(uint32_t c[2], uint32_t k0)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r) k0 += 0x9E3779B9u;
        const uint64_t p = (uint64_t)0xd256d193u * c[0];
        const uint32_t lo = (uint32_t)p, hi = (uint32_t)(p >> 32);
        c[0] = xor3_u32(hi, k0, c[1]);
        c[1] = lo;
    }
}

__device__ __forceinline__ void philox2x64_10(uint64_t c[2], uint64_t k0)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r) k0 += 0x9E3779B97F4A7C15ull;
        const uint64_t lo = 0xD2B74407B1CE6E93ull * c[0], hi = __umul64hi(0xD2B74407B1CE6E93ull, c[0]);
        c[0] = hi ^ k0 ^ c[1];
        c[1] = lo;
    }
}

template <int N>
__device__ __forceinline__ uint32_t rotl32(uint32_t x) { return (x << N) | (x >> (32 - N)); }
template <int N>
__device__ __forceinline__ uint64_t rotl64(uint64_t x) { return (x << N) | (x >> (64 - N)); }

// Chai's Threefry4x32-20: pairs (0,1),(2,3) every round; rotations R_32x4.
__device__ __forceinline__ void threefry4x32_20(uint32_t X[4], const uint32_t k[4])
{
    const uint32_t ks[5] = {k[0], k[1], k[2], k[3], 0x1BD11BDAu ^ k[0] ^ k[1] ^ k[2] ^ k[3]};
    X[0] += ks[0]; X[1] += ks[1]; X[2] += ks[2]; X[3] += ks[3];
#define CHAI_TF4_ROUND(R0, R1)                                   \
    X[0] += X[1]; X[1] = rotl32<R0>(X[1]); X[1] ^= X[0];         \
    X[2] += X[3]; X[3] = rotl32<R1>(X[3]); X[3] ^= X[2];
#define CHAI_TF4_INJECT(S)                                                   \
    X[0] += ks[(S) % 5]; X[1] += ks[((S) + 1) % 5]; X[2] += ks[((S) + 2) % 5]; \
    X[3] += ks[((S) + 3) % 5]; X[3] += (S);
    CHAI_TF4_ROUND(10, 26) CHAI_TF4_ROUND(11, 21) CHAI_TF4_ROUND(13, 27) CHAI_TF4_ROUND(23, 5)
    CHAI_TF4_INJECT(1)
    CHAI_TF4_ROUND(6, 20) CHAI_TF4_ROUND(17, 11) CHAI_TF4_ROUND(25, 10) CHAI_TF4_ROUND(18, 20)
    CHAI_TF4_INJECT(2)
    CHAI_TF4_ROUND(10, 26) CHAI_TF4_ROUND(11, 21) CHAI_TF4_ROUND(13, 27) CHAI_TF4_ROUND(23, 5)
    CHAI_TF4_INJECT(3)
    CHAI_TF4_ROUND(6, 20) CHAI_TF4_ROUND(17, 11) CHAI_TF4_ROUND(25, 10) CHAI_TF4_ROUND(18, 20)
    CHAI_TF4_INJECT(4)
    CHAI_TF4_ROUND(10, 26) CHAI_TF4_ROUND(11, 21) CHAI_TF4_ROUND(13, 27) CHAI_TF4_ROUND(23, 5)
    CHAI_TF4_INJECT(5)
#undef CHAI_TF4_ROUND
#undef CHAI_TF4_INJECT
}

__device__ __forceinline__ void threefry2x32_20(uint32_t X[2], uint32_t k0, uint32_t k1)
{
    const uint32_t ks[3] = {k0, k1, 0x1BD11BDAu ^ k0 ^ k1};
    X[0] += ks[0]; X[1] += ks[1];
#define CHAI_TF2_ROUND(R) X[0] += X[1]; X[1] = rotl32<R>(X[1]); X[1] ^= X[0];
#define CHAI_TF2_INJECT(S) X[0] += ks[(S) % 3]; X[1] += ks[((S) + 1) % 3]; X[1] += (S);
    CHAI_TF2_ROUND(13) CHAI_TF2_ROUND(15) CHAI_TF2_ROUND(26) CHAI_TF2_ROUND(6) CHAI_TF2_INJECT(1)
    CHAI_TF2_ROUND(17) CHAI_TF2_ROUND(29) CHAI_TF2_ROUND(16) CHAI_TF2_ROUND(24) CHAI_TF2_INJECT(2)
    CHAI_TF2_ROUND(13) CHAI_TF2_ROUND(15) CHAI_TF2_ROUND(26) CHAI_TF2_ROUND(6) CHAI_TF2_INJECT(3)
    CHAI_TF2_ROUND(17) CHAI_TF2_ROUND(29) CHAI_TF2_ROUND(16) CHAI_TF2_ROUND(24) CHAI_TF2_INJECT(4)
    CHAI_TF2_ROUND(13) CHAI_TF2_ROUND(15) CHAI_TF2_ROUND(26) CHAI_TF2_ROUND(6) CHAI_TF2_INJECT(5)
#undef CHAI_TF2_ROUND
#undef CHAI_TF2_INJECT
}

__device__ __forceinline__ void threefry2x64_20(uint64_t X[2], uint64_t k0, uint64_t k1)
{
    const uint64_t ks[3] = {k0, k1, 0x1BD11BDAA9FC1A22ull ^ k0 ^ k1};
    X[0] += ks[0]; X[1] += ks[1];
#define CHAI_TF2_ROUND(R) X[0] += X[1]; X[1] = rotl64<R>(X[1]); X[1] ^= X[0];
#define CHAI_TF2_INJECT(S) X[0] += ks[(S) % 3]; X[1] += ks[((S) + 1) % 3]; X[1] += (S);
    CHAI_TF2_ROUND(16) CHAI_TF2_ROUND(42) CHAI_TF2_ROUND(12) CHAI_TF2_ROUND(31) CHAI_TF2_INJECT(1)
    CHAI_TF2_ROUND(16) CHAI_TF2_ROUND(32) CHAI_TF2_ROUND(24) CHAI_TF2_ROUND(21) CHAI_TF2_INJECT(2)
    CHAI_TF2_ROUND(16) CHAI_TF2_ROUND(42) CHAI_TF2_ROUND(12) CHAI_TF2_ROUND(31) CHAI_TF2_INJECT(3)
    CHAI_TF2_ROUND(16) CHAI_TF2_ROUND(32) CHAI_TF2_ROUND(24) CHAI_TF2_ROUND(21) CHAI_TF2_INJECT(4)
    CHAI_TF2_ROUND(16) CHAI_TF2_ROUND(42) CHAI_TF2_ROUND(12) CHAI_TF2_ROUND(31) CHAI_TF2_INJECT(5)
#undef CHAI_TF2_ROUND
#undef CHAI_TF2_INJECT
}

// ------------------------------------------------------- uniform conversions
// (float)(int32)w * 2^-32 + 0.5 as one fma: the product by a power of two is
// exact (|w| >= 1 stays normal), so the fused form rounds once, where the
// two-step form rounds only at the add -- the same value
__device__ __forceinline__ float unit_f32(uint32_t w)
{
    return __builtin_fmaf((float)(int32_t)w, 0x1p-32f, 0.5f);
}

__device__ __forceinline__ double unit_f64(uint64_t w)
{
    return __builtin_fma((double)(int64_t)w, 0x1p-64, 0.5);
}

// ------------------------------------------------------ Box-Muller math (f64)
// log(x) = e*ln2 + 2 atanh(s), s = f/(2+f), m = 1+f in [sqrt(1/2), sqrt(2));
// 2 atanh(s) = 2s + s z P(z), P(z) = sum_{k=1..10} 2/(2k+1) z^(k-1), z = s^2.
__device__ __forceinline__ double bm_log(double x)
{
    if (!(x > 0.0)) return x == 0.0 ? -__builtin_inf() : __builtin_nan("");
    if (__builtin_isinf(x)) return x;
    uint64_t b = __double_as_longlong(x);
    int e = (int)((b >> 52) & 0x7ff);
    if (e == 0) {
        x = x * 0x1p54;
        b = __double_as_longlong(x);
        e = (int)((b >> 52) & 0x7ff) - 54;
    }
    e -= 1023;
    double m = __longlong_as_double((long long)((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull));
    if (m > 0x1.6a09e667f3bcdp+0) { m = m * 0.5; e += 1; }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double p = 0x1.8618618618618p-4;
    p = __builtin_fma(p, z, 0x1.af286bca1af28p-4);
    p = __builtin_fma(p, z, 0x1.e1e1e1e1e1e1ep-4);
    p = __builtin_fma(p, z, 0x1.1111111111111p-3);
    p = __builtin_fma(p, z, 0x1.3b13b13b13b14p-3);
    p = __builtin_fma(p, z, 0x1.745d1745d1746p-3);
    p = __builtin_fma(p, z, 0x1.c71c71c71c71cp-3);
    p = __builtin_fma(p, z, 0x1.2492492492492p-2);
    p = __builtin_fma(p, z, 0x1.999999999999ap-2);
    p = __builtin_fma(p, z, 0x1.5555555555555p-1);
    const double sz = s * z;
    const double logm = __builtin_fma(sz, p, 2.0 * s);
    const double de = (double)e;
    return __builtin_fma(de, 0x1.62e42fee00000p-1, __builtin_fma(de, 0x1.a39ef35793c76p-33, logm));
}

// sin/cos(2 pi u): q = rint(4u), r = u - q/4 (exact), x = 2 pi r, Taylor series.
__device__ __forceinline__ void bm_sincos2pi(double u, double& s, double& c)
{
    const double q = __builtin_rint(4.0 * u);
    const double r = u - 0.25 * q;
    const int qi = ((int)(long long)q) & 3;
    const double x = r * 0x1.921fb54442d18p+2;
    const double x2 = x * x;
    double ps = 0x1.952c77030ad4ap-49;
    ps = __builtin_fma(ps, x2, -0x1.ae7f3e733b81fp-41);
    ps = __builtin_fma(ps, x2, 0x1.6124613a86d09p-33);
    ps = __builtin_fma(ps, x2, -0x1.ae64567f544e4p-26);
    ps = __builtin_fma(ps, x2, 0x1.71de3a556c734p-19);
    ps = __builtin_fma(ps, x2, -0x1.a01a01a01a01ap-13);
    ps = __builtin_fma(ps, x2, 0x1.1111111111111p-7);
    ps = __builtin_fma(ps, x2, -0x1.5555555555555p-3);
    const double sn = __builtin_fma(x * x2, ps, x);
    double pc = -0x1.6827863b97d97p-53;
    pc = __builtin_fma(pc, x2, 0x1.ae7f3e733b81fp-45);
    pc = __builtin_fma(pc, x2, -0x1.93974a8c07c9dp-37);
    pc = __builtin_fma(pc, x2, 0x1.1eed8eff8d898p-29);
    pc = __builtin_fma(pc, x2, -0x1.27e4fb7789f5cp-22);
    pc = __builtin_fma(pc, x2, 0x1.a01a01a01a01ap-16);
    pc = __builtin_fma(pc, x2, -0x1.6c16c16c16c17p-10);
    pc = __builtin_fma(pc, x2, 0x1.5555555555555p-5);
    pc = __builtin_fma(pc, x2, -0.5);
    const double cs = __builtin_fma(x2, pc, 1.0);
    switch (qi) {
        case 0: s = sn; c = cs; break;
        case 1: s = cs; c = -sn; break;
        case 2: s = -sn; c = -cs; break;
        default: s = -cs; c = sn; break;
    }
}

__device__ __forceinline__ void box_muller(double ua, double ub, double& n0, double& n1)
{
    const double r = __builtin_sqrt(-2.0 * bm_log(ua));
    double s, c;
    bm_sincos2pi(ub, s, c);
    n0 = r * s;
    n1 = r * c;
}

// ------------------------------------------------------ Box-Muller math (f32)
// The f32 stream evaluates its normals in f32: the same range reductions as
// above with series truncated for single precision, every operation an IEEE
// f32 op (explicit fma, correctly rounded divide and sqrt), so the host
// restatement reproduces them bit for bit.
// log(x) = e*ln2 + 2s + s z P(z), s = f/(2+f), z = s^2, P(z) = sum_{k=1..5} 2/(2k+1) z^(k-1).
__device__ __forceinline__ float bm_logf(float x)
{
    if (!(x > 0.0f)) return x == 0.0f ? -__builtin_inff() : __builtin_nanf("");
    if (__builtin_isinf(x)) return x;
    uint32_t b = __float_as_uint(x);
    int e = (int)((b >> 23) & 0xff);
    if (e == 0) {
        x = x * 0x1p25f;
        b = __float_as_uint(x);
        e = (int)((b >> 23) & 0xff) - 25;
    }
    e -= 127;
    float m = __uint_as_float((b & 0x007fffffu) | 0x3f800000u);
    if (m > 0x1.6a09e6p+0f) { m = m * 0.5f; e += 1; }
    const float f = m - 1.0f;
    const float s = f / (2.0f + f);
    const float z = s * s;
    float p = 0x1.745d18p-3f;
    p = __builtin_fmaf(p, z, 0x1.c71c72p-3f);
    p = __builtin_fmaf(p, z, 0x1.24924ap-2f);
    p = __builtin_fmaf(p, z, 0x1.99999ap-2f);
    p = __builtin_fmaf(p, z, 0x1.555556p-1f);
    const float logm = __builtin_fmaf(s * z, p, 2.0f * s);
    const float de = (float)e;
    return __builtin_fmaf(de, 0x1.62e4p-1f, __builtin_fmaf(de, 0x1.7f7d1cp-20f, logm));
}

// sin/cos(2 pi u): q = rint(4u), t = u - q/4 in [-1/8, 1/8] (exact); series in t
// with coefficients (2 pi)^k / k! through t^9 (sin) and t^10 (cos).
__device__ __forceinline__ void bm_sincos2pif(float u, float& s, float& c)
{
    const float q = __builtin_rintf(4.0f * u);
    const float t = u - 0.25f * q;
    const int qi = ((int)q) & 3;
    const float t2 = t * t;
    float ps = 0x1.507834p+5f;
    ps = __builtin_fmaf(ps, t2, -0x1.32d2ccp+6f);
    ps = __builtin_fmaf(ps, t2, 0x1.466bc6p+6f);
    ps = __builtin_fmaf(ps, t2, -0x1.4abbcep+5f);
    const float sn = __builtin_fmaf(t * t2, ps, t * 0x1.921fb6p+2f);
    float pc = -0x1.a6d1f2p+4f;
    pc = __builtin_fmaf(pc, t2, 0x1.e1f506p+5f);
    pc = __builtin_fmaf(pc, t2, -0x1.55d3c8p+6f);
    pc = __builtin_fmaf(pc, t2, 0x1.03c1f0p+6f);
    pc = __builtin_fmaf(pc, t2, -0x1.3bd3ccp+4f);
    const float cs = __builtin_fmaf(t2, pc, 1.0f);
    // quadrant qi: (s, c) = (sn, cs), (cs, -sn), (-sn, -cs), (-cs, sn) -- a swap
    // for odd qi and sign flips, as selects and xors rather than branches
    const bool odd = qi & 1;
    s = __uint_as_float(__float_as_uint(odd ? cs : sn) ^ ((uint32_t)(qi & 2) << 30));
    c = __uint_as_float(__float_as_uint(odd ? sn : cs) ^ ((uint32_t)((qi + 1) & 2) << 30));
}

// The f32 normals' own forms of log and sqrt.  Their inputs come from
// unit_f32, which gives 0, 1 or a normal number in between, so these skip
// the general routines' scaling and special cases and their control flow,
// and replace the expanded IEEE divide and square root with short
// sequences.  Each returns the bits of the general form (bm_logf with its
// IEEE divide, the correctly rounded sqrtf) for every one of the 2^32 stream
// words, checked exhaustively on the device.

// f / d for d = 2 + f in [1.70, 2.42] and f = 0 or 2^-23 <= |f| < 0.42: the
// hardware reciprocal refined by one Newton step, the quotient corrected once
// by its residual (exact in the fma).
__device__ __forceinline__ float bm_div(float f, float d)
{
    float r = __builtin_amdgcn_rcpf(d);
    r = __builtin_fmaf(__builtin_fmaf(-d, r, 1.0f), r, r);
    const float q = f * r;
    return __builtin_fmaf(__builtin_fmaf(-d, q, f), r, q);
}

// bm_logf for x = 0, 1 or a normal number between.
__device__ __forceinline__ float bm_logf_unit(float x)
{
    const uint32_t b = __float_as_uint(x);
    float m = __uint_as_float((b & 0x007fffffu) | 0x3f800000u);
    const bool up = m > 0x1.6a09e6p+0f;
    m = up ? m * 0.5f : m;
    const int e = (int)(b >> 23) - 127 + (up ? 1 : 0);
    const float f = m - 1.0f;
    const float s = bm_div(f, 2.0f + f);
    const float z = s * s;
    float p = 0x1.745d18p-3f;
    p = __builtin_fmaf(p, z, 0x1.c71c72p-3f);
    p = __builtin_fmaf(p, z, 0x1.24924ap-2f);
    p = __builtin_fmaf(p, z, 0x1.99999ap-2f);
    p = __builtin_fmaf(p, z, 0x1.555556p-1f);
    const float logm = __builtin_fmaf(s * z, p, 2.0f * s);
    const float de = (float)e;
    const float l = __builtin_fmaf(de, 0x1.62e4p-1f, __builtin_fmaf(de, 0x1.7f7d1cp-20f, logm));
    return x == 0.0f ? -__builtin_inff() : l;
}

// Correctly rounded sqrt of y = -2 log(u): -0, +inf or a normal number.  The
// hardware square root s lies within an ulp of sqrt(y); y <= s_dn * s (the
// square of the midpoint below s, less ulp^2 / 4) picks the neighbour below,
// y > s_up * s the one above.
__device__ __forceinline__ float bm_sqrtf(float y)
{
    const float s = __builtin_amdgcn_sqrtf(y);
    const float sd = __uint_as_float(__float_as_uint(s) - 1u), su = __uint_as_float(__float_as_uint(s) + 1u);
    float t = __builtin_fmaf(-sd, s, y) <= 0.0f ? sd : s;
    t = __builtin_fmaf(-su, s, y) > 0.0f ? su : t;
    return y == 0.0f || y == __builtin_inff() ? y : t;
}

__device__ __forceinline__ void box_muller_f32(float ua, float ub, float& n0, float& n1)
{
    const float r = bm_sqrtf(-2.0f * bm_logf_unit(ua));
    float s, c;
    bm_sincos2pif(ub, s, c);
    n0 = r * s;
    n1 = r * c;
}

// ---------------------------------------------------------- stream vectors
// 32-bit words of one generated vector (4 words for vl 4, 2 for vl 1/2).
template <int VARIANT, int VL>
__device__ __forceinline__ void rng_words32(uint32_t vidx, uint32_t seed, uint32_t w[4])
{
    if (VL == 4) {
        w[0] = seed; w[1] = 0xf00dcafeu; w[2] = 0xdeadbeefu; w[3] = 0xbeeff00du;
        if (VARIANT == 1231) {
            const uint32_t k[4] = {vidx, 0xdecafbadu, 0xfacebeadu, 0x12345678u};
            threefry4x32_20(w, k);
        } else {
            philox4x32_10(w, vidx, 0xdecafbadu);
        }
    } else {
        w[0] = seed; w[1] = 0xf00dcafeu; w[2] = 0; w[3] = 0;
        if (VARIANT == 1231) threefry2x32_20(w, vidx, 0xdecafbadu);
        else philox2x32_10(w, vidx);
    }
}

template <int VARIANT>
__device__ __forceinline__ void rng_words64(uint64_t vidx, uint64_t seed, uint64_t w[2])
{
    w[0] = seed; w[1] = 0xf00dcafeull;
    if (VARIANT == 1231) threefry2x64_20(w, vidx, 0xdecafbadull);
    else philox2x64_10(w, vidx);
}

// f32 normal vector: v[0..3] from the 4 (or 2) words.
__device__ __forceinline__ void normals_f32(const uint32_t w[4], int pairs, float v[4])
{
    for (int p = 0; p < pairs; p++) box_muller_f32(unit_f32(w[2 * p]), unit_f32(w[2 * p + 1]), v[2 * p], v[2 * p + 1]);
}

}  // namespace chai_hip
