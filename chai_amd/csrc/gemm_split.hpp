// gemm_split.hpp -- the split-bf16 pre-pass tile shared by the dense and the
// ragged GEMM paths.
//
// op(X) element (r, k) lives at src[r * ld + k] (TRANS = false) or
// src[k * ld + r] (TRANS = true).  The output panels hi / lo hold Rp rows and
// Kp columns of bf16 in k blocks of 32: element (r, k) at
// (k / 32) * Rp * 32 + r * 32 + k % 32, so that the 32-deep slice of a row
// range is one contiguous run (the GEMM loads it straight into LDS).  X = hi +
// lo (+ a residual of at most 2^-16 |X|), with zeros wherever r >= R or k >= K.
// One call covers the 64 x 64 block at (r0, k0) with 256 threads; reads are
// 16-byte vectors along the contiguous source index.  A K-contiguous source is
// split straight from registers; a transposed one goes through a padded LDS
// tile so that each thread can write 16 consecutive k.
#pragma once
#include "gemm_device.hpp"

namespace chai_hip {

template <typename T>
__device__ __forceinline__ float to_f32(T v) { return (float)v; }

template <typename T>
__device__ __forceinline__ f32x4 load4_f32(const T* p)
{
    if constexpr (sizeof(T) == 4 && (T)0.5 != 0) {
        return gload(reinterpret_cast<const f32x4*>(p));
    } else {
        typedef T t4 __attribute__((ext_vector_type(4)));
        const t4 v = gload(reinterpret_cast<const t4*>(p));
        return f32x4{(float)v[0], (float)v[1], (float)v[2], (float)v[3]};
    }
}

// The third bf16 piece of two values: lo2 = bf16(x - hi - lo).  x - hi is
// exact (Sterbenz, as in split_bf16x2), and so is (x - hi) - lo: lo is the
// nearest bf16 to x - hi, within a factor two of it.  hi + lo + lo2 carries 24
// significant bits of x, all of an f32.
__device__ __forceinline__ uint32_t split_third_bf16x2(float a0, float a1, uint32_t hi, uint32_t lo)
{
    const f32x2_t v = {a0, a1};
    const f32x2_t hf = {__uint_as_float(hi << 16), __uint_as_float(hi & 0xffff0000u)};
    const f32x2_t lf = {__uint_as_float(lo << 16), __uint_as_float(lo & 0xffff0000u)};
    const f32x2_t r = (v - hf) - lf;
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(r, bf16x2_t));
}

// 16 consecutive k of one row -> 2 x 16 bytes of each panel (hi, lo and, when
// THREE, lo2); returns the OR of the lo words stored (zero iff every value
// stored was exact in bf16)
template <bool THREE = false>
__device__ __forceinline__ uint32_t split_store16(const float* v, uint16_t* hi, uint16_t* lo, uint64_t off,
                                                 bool second, uint16_t* lo2 = nullptr)
{
    u32x4 vh0, vh1, vl0, vl1, vm0, vm1;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        uint32_t h, l;
        split_bf16x2(v[2 * j], v[2 * j + 1], h, l);
        vh0[j] = h; vl0[j] = l;
        if (THREE) vm0[j] = split_third_bf16x2(v[2 * j], v[2 * j + 1], h, l);
        split_bf16x2(v[8 + 2 * j], v[8 + 2 * j + 1], h, l);
        vh1[j] = h; vl1[j] = l;
        if (THREE) vm1[j] = split_third_bf16x2(v[8 + 2 * j], v[8 + 2 * j + 1], h, l);
    }
    u32x4* dh = reinterpret_cast<u32x4*>(hi + off);
    u32x4* dl = reinterpret_cast<u32x4*>(lo + off);
    gstore(dh, vh0);
    gstore(dl, vl0);
    if (THREE) gstore(reinterpret_cast<u32x4*>(lo2 + off), vm0);
    uint32_t lo_or = vl0[0] | vl0[1] | vl0[2] | vl0[3];
    if (second) {
        gstore(dh + 1, vh1); gstore(dl + 1, vl1);
        if (THREE) gstore(reinterpret_cast<u32x4*>(lo2 + off) + 1, vm1);
        lo_or |= vl1[0] | vl1[1] | vl1[2] | vl1[3];
    }
    return lo_or;
}

// offset of panel element (r, k), k a multiple of 16, in the k-blocked layout
__device__ __forceinline__ uint64_t panel_off(uint64_t r, uint64_t k, uint64_t Rp)
{
    return (k >> 5) * (Rp * 32) + r * 32 + (k & 31);
}

// Returns the OR of this thread's lo words (see split_store16).  THREE also
// writes the third piece into lo2.
template <typename T, bool TRANS, bool THREE = false>
__device__ __forceinline__ uint32_t split_tile(const T* __restrict__ src, uint64_t ld, uint64_t R,
                                           uint64_t K, uint64_t Kp, uint64_t Rp,
                                           uint16_t* __restrict__ hi, uint16_t* __restrict__ lo,
                                           uint64_t r0, uint64_t k0, float (*tile)[65],
                                           uint16_t* __restrict__ lo2 = nullptr)
{
    const int t = threadIdx.x;
    // each thread finally writes 16 consecutive k of one row
    const int rr = t >> 2, kq = (t & 3) * 16;
    const uint64_t r = r0 + rr;
    const bool vec = ((ld & 3) == 0) && (((uintptr_t)src & 15) == 0);
    if (!TRANS) {
        // K-contiguous source: no transpose needed, straight from registers
        if (k0 + kq >= Kp) return 0u;
        float v[16];
        const uint64_t kb = k0 + kq;
        if (vec && r < R && kb + 16 <= K) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const f32x4 x = load4_f32(src + r * ld + kb + 4 * q);
                v[4 * q] = x[0]; v[4 * q + 1] = x[1]; v[4 * q + 2] = x[2]; v[4 * q + 3] = x[3];
            }
        } else {
#pragma unroll
            for (int j = 0; j < 16; j++) v[j] = (r < R && kb + j < K) ? to_f32(gload(src + r * ld + kb + j)) : 0.0f;
        }
        return split_store16<THREE>(v, hi, lo, panel_off(r, kb, Rp), kb + 8 < Kp, lo2);
    }
    // source element (r, k) at src[k * ld + r]: read 4 consecutive r per lane
    // (16 lanes per 64-wide row), transpose through the padded LDS tile
    {
        const int r4 = (t & 15) * 4, kk0 = t >> 4;   // 16 k rows per pass
        const bool fullr = vec && r0 + r4 + 4 <= R;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int kk = kk0 + 16 * i;
            const uint64_t k = k0 + kk;
            if (fullr && k < K) {
                const f32x4 x = load4_f32(src + k * ld + r0 + r4);
                tile[kk][r4] = x[0]; tile[kk][r4 + 1] = x[1]; tile[kk][r4 + 2] = x[2]; tile[kk][r4 + 3] = x[3];
            } else {
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const uint64_t rj = r0 + r4 + j;
                    tile[kk][r4 + j] = (rj < R && k < K) ? to_f32(gload(src + k * ld + rj)) : 0.0f;
                }
            }
        }
    }
    __syncthreads();
    if (k0 + kq >= Kp) return 0u;
    float v[16];
#pragma unroll
    for (int j = 0; j < 16; j++) v[j] = tile[kq + j][rr];
    return split_store16<THREE>(v, hi, lo, panel_off(r, k0 + kq, Rp), k0 + kq + 8 < Kp, lo2);
}

}  // namespace chai_hip
