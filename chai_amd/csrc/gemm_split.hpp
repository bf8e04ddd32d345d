// gemm_split.hpp -- the split-bf16 pre-pass tile shared by the dense and the
// ragged GEMM paths.
//
// op(X) element (r, k) lives at src[r * ld + k] (TRANS = false) or
// src[k * ld + r] (TRANS = true).  The output panels hi / lo are [Rp][Kp]
// bf16, K contiguous, with X = hi + lo (+ a residual below 2^-17 |X|) and zeros
// wherever r >= R or k >= K.  One call covers the 64 x 64 block at (r0, k0)
// with 256 threads; reads are coalesced along the contiguous source index and
// go through a padded LDS tile so that each thread can write 16 consecutive k.
#pragma once
#include "gemm_device.hpp"

namespace chai_hip {

template <typename T>
__device__ __forceinline__ float to_f32(T v) { return (float)v; }

template <typename T, bool TRANS>
__device__ __forceinline__ void split_tile(const T* __restrict__ src, uint64_t ld, uint64_t R,
                                           uint64_t K, uint64_t Kp, uint16_t* __restrict__ hi,
                                           uint16_t* __restrict__ lo, uint64_t r0, uint64_t k0,
                                           float (*tile)[65])
{
    const int t = threadIdx.x;
    if (TRANS) {
        // 64 lanes read 64 consecutive r at one k
        for (int i = 0; i < 16; i++) {
            const int kk = i * 4 + (t >> 6), rr = t & 63;
            const uint64_t r = r0 + rr, k = k0 + kk;
            tile[kk][rr] = (r < R && k < K) ? to_f32(src[k * ld + r]) : 0.0f;
        }
    } else {
        for (int i = 0; i < 16; i++) {
            const int rr = i * 4 + (t >> 6), kk = t & 63;
            const uint64_t r = r0 + rr, k = k0 + kk;
            tile[kk][rr] = (r < R && k < K) ? to_f32(src[r * ld + k]) : 0.0f;
        }
    }
    __syncthreads();
    // each thread writes 16 consecutive k of one row: 2 x 16 bytes per panel
    const int rr = t >> 2, kq = (t & 3) * 16;
    const uint64_t r = r0 + rr;
    uint16_t h[16], l[16];
#pragma unroll
    for (int j = 0; j < 16; j++) split_bf16(tile[kq + j][rr], h[j], l[j]);
    if (k0 + kq < Kp) {
        u32x4* dh = reinterpret_cast<u32x4*>(hi + r * Kp + k0 + kq);
        u32x4* dl = reinterpret_cast<u32x4*>(lo + r * Kp + k0 + kq);
        u32x4 vh0, vh1, vl0, vl1;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            vh0[j] = h[2 * j] | ((uint32_t)h[2 * j + 1] << 16);
            vh1[j] = h[8 + 2 * j] | ((uint32_t)h[8 + 2 * j + 1] << 16);
            vl0[j] = l[2 * j] | ((uint32_t)l[2 * j + 1] << 16);
            vl1[j] = l[8 + 2 * j] | ((uint32_t)l[8 + 2 * j + 1] << 16);
        }
        dh[0] = vh0;
        if (k0 + kq + 8 < Kp) dh[1] = vh1;
        dl[0] = vl0;
        if (k0 + kq + 8 < Kp) dl[1] = vl1;
    }
}

}  // namespace chai_hip
