// gemm_x3_f32.hpp -- the split-BF16x3 tile fed straight from f32 operands.
//
// Row-major A (M x K, row stride lda) and B (K x N, row stride ldb) are read
// into registers one 32-deep stage ahead, split into bf16 hi / lo there (the
// same round-to-nearest-even conversion the split pre-pass uses) and written
// to the swizzled LDS layout of the panel tile (gemm_x3.hpp).  The MFMA
// sequence per output element is the panel tile's, so results are equal bit
// for bit; what goes away is the pre-pass and its panel workspace, at the
// price of conversion work and register staging inside the tile.  Rows,
// columns and k beyond M, N, K read as zeros; the stage count is Kp / 32 with
// Kp = K rounded up to 64, as for the panels.
//
// Load mapping (T threads, BK = 32): two threads per A row, each 16
// consecutive k (four 16-byte loads when aligned); two threads per B column,
// each 16 consecutive k of it (a wave's scalar loads cover 64 consecutive
// columns of one k row).  Each thread then writes two 16-byte chunks per
// panel: 8 consecutive k of one row, the layout the MFMA fragments read.
#pragma once
#include "gemm_x3.hpp"

namespace chai_hip {

template <int BM, int BN, int WGM, int WGN>
struct X3F32Shape {
    using S = X3Shape<BM, BN, 32, WGM, WGN>;
    static constexpr int T = S::THREADS;
    static constexpr int RP = T / 2;           // A rows / B columns per pass
    static constexpr int NPA = BM / RP, NPB = BN / RP;
    static_assert(BM % RP == 0 && BN % RP == 0, "whole passes");
};

// 16 consecutive k of one row (zeros outside the operand) -> hi / lo chunk
// pairs at LDS chunk offsets o0 (k 0..7) and o1 (k 8..15) of the hi panel;
// the lo panel is `plane` chunks further.
__device__ __forceinline__ void x3f_write16(const float* v, u32x4* lds, int o0, int o1, int plane)
{
    u32x4 h0, h1, l0, l1;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        uint32_t h, l;
        split_bf16x2(v[2 * j], v[2 * j + 1], h, l);
        h0[j] = h; l0[j] = l;
        split_bf16x2(v[8 + 2 * j], v[8 + 2 * j + 1], h, l);
        h1[j] = h; l1[j] = l;
    }
    lds[o0] = h0; lds[o0 + plane] = l0;
    lds[o1] = h1; lds[o1 + plane] = l1;
}

template <int BM, int BN, int WGM, int WGN>
__device__ __forceinline__ void x3f_load(const GemmArgs& a, uint64_t r0, uint64_t c0, uint64_t k0, bool avec,
                                         float (&fa)[X3F32Shape<BM, BN, WGM, WGN>::NPA][16],
                                         float (&fb)[X3F32Shape<BM, BN, WGM, WGN>::NPB][16])
{
    using F = X3F32Shape<BM, BN, WGM, WGN>;
    const int t = threadIdx.x;
    // A stage whose 32 k all lie below K (every stage but the last of a
    // K % 32 != 0 operand) needs one row / column test per lane instead of
    // one per element; kfull is block-uniform.
    const bool kfull = k0 + 32 <= a.K;
    const uint64_t kh = k0 + (uint64_t)(t & 1) * 16;          // A: 16 k of row t / 2
#pragma unroll
    for (int p = 0; p < F::NPA; p++) {
        const uint64_t row = r0 + (uint64_t)(t >> 1) + (uint64_t)p * F::RP;
        const float* src = a.Af + row * a.lda + kh;
        if (kfull && avec) {
            if (row < a.M) {
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const f32x4 x = gload(reinterpret_cast<const f32x4*>(src) + q);
                    fa[p][4 * q] = x[0]; fa[p][4 * q + 1] = x[1]; fa[p][4 * q + 2] = x[2]; fa[p][4 * q + 3] = x[3];
                }
            } else {
#pragma unroll
                for (int j = 0; j < 16; j++) fa[p][j] = 0.0f;
            }
        } else {
#pragma unroll
            for (int j = 0; j < 16; j++) fa[p][j] = (row < a.M && kh + j < a.K) ? gload(src + j) : 0.0f;
        }
    }
    const uint64_t kb = k0 + (uint64_t)(t / F::RP) * 16;       // B: 16 k of column t % RP
#pragma unroll
    for (int p = 0; p < F::NPB; p++) {
        const uint64_t col = c0 + (uint64_t)(t % F::RP) + (uint64_t)p * F::RP;
        const float* src = a.Bf + kb * a.ldb + col;
        if (kfull) {
            if (col < a.N) {
#pragma unroll
                for (int j = 0; j < 16; j++) fb[p][j] = gload(src + (uint64_t)j * a.ldb);
            } else {
#pragma unroll
                for (int j = 0; j < 16; j++) fb[p][j] = 0.0f;
            }
        } else {
#pragma unroll
            for (int j = 0; j < 16; j++)
                fb[p][j] = (col < a.N && kb + j < a.K) ? gload(src + (uint64_t)j * a.ldb) : 0.0f;
        }
    }
}

template <int BM, int BN, int WGM, int WGN>
__device__ __forceinline__ void x3f_store_stage(u32x4* st, const float (&fa)[X3F32Shape<BM, BN, WGM, WGN>::NPA][16],
                                                const float (&fb)[X3F32Shape<BM, BN, WGM, WGN>::NPB][16])
{
    using F = X3F32Shape<BM, BN, WGM, WGN>;
    using S = typename F::S;
    const int t = threadIdx.x;
    const int qa = (t & 1) * 2, qb = (t / F::RP) * 2;          // first 8-k chunk of each thread's 16 k
#pragma unroll
    for (int p = 0; p < F::NPA; p++) {
        const int row = (t >> 1) + p * F::RP;
        const int sw = (row >> S::SH) & (S::CPR - 1);
        x3f_write16(fa[p], st, row * S::CPR + (qa ^ sw), row * S::CPR + ((qa + 1) ^ sw), S::PA);
    }
#pragma unroll
    for (int p = 0; p < F::NPB; p++) {
        const int row = (t % F::RP) + p * F::RP;
        const int sw = (row >> S::SH) & (S::CPR - 1);
        x3f_write16(fb[p], st, 2 * S::PA + row * S::CPR + (qb ^ sw), 2 * S::PA + row * S::CPR + ((qb + 1) ^ sw),
                    S::PB);
    }
}

// One BM x BN output tile (tm, tn) from f32 A (a.Af, a.lda) and B (a.Bf,
// a.ldb); a.Kp = K rounded up to 64.  LDS is free on entry and on exit.
// PIPE: software-pipeline the MFMA operand fragments across the barrier.
template <int BM, int BN, int WGM, int WGN, bool PIPE>
__device__ __forceinline__ void x3f_tile(const GemmArgs& a, uint32_t tm, uint32_t tn, u32x4* lds)
{
    using F = X3F32Shape<BM, BN, WGM, WGN>;
    using S = typename F::S;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = (wave / WGN) * S::WM, wn = (wave % WGN) * S::WN;
    const uint64_t r0 = (uint64_t)tm * BM, c0 = (uint64_t)tn * BN;
    const bool avec = ((a.lda & 3) == 0) && (((uintptr_t)a.Af & 15) == 0);

    floatx16 acc[S::TI][S::TJ];
#pragma unroll
    for (int i = 0; i < S::TI; i++)
#pragma unroll
        for (int j = 0; j < S::TJ; j++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[i][j][r] = 0.0f;

    float fa[F::NPA][16], fb[F::NPB][16];
    const int nk = (int)(a.Kp / 32);
    x3f_load<BM, BN, WGM, WGN>(a, r0, c0, 0, avec, fa, fb);
    x3f_store_stage<BM, BN, WGM, WGN>(lds, fa, fb);
    __syncthreads();

    int offA[S::TI][S::KB], offB[S::TJ][S::KB];
    x3_frag_offsets<S>(wm, wn, lane, offA, offB);

    if constexpr (!PIPE) {
        for (int kt = 0; kt < nk; kt++) {
            const int cur = (kt & 1) * S::STAGE;
            const bool more = kt + 1 < nk;
            if (more) x3f_load<BM, BN, WGM, WGN>(a, r0, c0, (uint64_t)(kt + 1) * 32, avec, fa, fb);
            x3_stage_mfma<S, 0>(lds, cur, offA, offB, acc);
            // the other buffer was last read in step kt - 1, closed by its barrier
            if (more) x3f_store_stage<BM, BN, WGM, WGN>(lds + ((kt + 1) & 1) * S::STAGE, fa, fb);
            __syncthreads();
        }
        x3_store<S>(a, a.C, r0 + wm, c0 + wn, lane, acc);
        return;
    }
    // The panel tile's fragment pipeline (gemm_x3.hpp): k block 1 of stage kt
    // is read before the barrier and multiplied after it, beside the reads of
    // stage kt + 1's k block 0.  It holds two k blocks of fragments besides
    // the staged operands, so it is used where a block's waves have a SIMD's
    // registers to themselves.
    static_assert(S::KB == 2, "32-deep stages");
    X3Frag<S> f[2];
    x3_load_frag<S, 0>(lds, 0, 0, offA, offB, f[0]);
    for (int kt = 0; kt < nk; kt++) {
        const int cur = (kt & 1) * S::STAGE;
        const bool more = kt + 1 < nk;
        if (more) x3f_load<BM, BN, WGM, WGN>(a, r0, c0, (uint64_t)(kt + 1) * 32, avec, fa, fb);
        x3_load_frag<S, 0>(lds, cur, 1, offA, offB, f[1]);
        x3_frag_mfma<S, 0>(f[0], acc);
        __builtin_amdgcn_sched_barrier(0);
        // the other buffer was last read before the previous step's barrier
        if (more) x3f_store_stage<BM, BN, WGM, WGN>(lds + ((kt + 1) & 1) * S::STAGE, fa, fb);
        __syncthreads();
        __builtin_amdgcn_sched_barrier(0);
        if (more) x3_load_frag<S, 0>(lds, S::STAGE - cur, 0, offA, offB, f[0]);
        x3_frag_mfma<S, 0>(f[1], acc);
    }
    x3_store<S>(a, a.C, r0 + wm, c0 + wn, lane, acc);
}

// The f32-fed family's tile shapes: X(BM, BN, waves_m, waves_n, min blocks per CU)
#define CHAI_X3F_VARIANTS(X)                                                  \
    X(128, 128, 2, 2, 2)   /* 4 waves of 64x64, 2 blocks / CU */             \
    X(256, 128, 2, 2, 1)   /* 4 waves of 128x64                */

}  // namespace chai_hip
