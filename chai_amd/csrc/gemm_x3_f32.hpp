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
#include <type_traits>
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

// Stage k0 of the tile's operands into registers: thread t holds 16
// consecutive k of A row t / 2 (+ p RP) and 16 consecutive k of B column
// t % RP (+ p RP).  The loads go through raw buffer descriptors, which are
// wave-uniform: A's over the 32 rows a wave reads ((rows - 1) lda + K floats
// from its first row), B's over the stage's k rows ((rows - 1) ldb + N floats
// from row k0).  A load past a descriptor's extent returns zero and touches
// no memory, which zeroes rows of A past M and k past K; each lane computes
// one offset per A row / B column and B's k rows are scalar offsets, so there
// is no per-element address arithmetic.  Columns of B past N are read at
// column N - 1 and zeroed.  Extents and offsets stay below 2^31 bytes for
// row strides below 2^23 (the ragged plan checks it).  FULL: the stage's 32 k
// lie below K (no k tests); AVEC: A's rows allow 16-byte loads (lda a
// multiple of 4, A 16-byte aligned).  Both are template arguments so that a
// loop of full stages is straight-line code whose load waits the compiler
// can count.
template <int BM, int BN, int WGM, int WGN, bool FULL, bool AVEC>
__device__ __forceinline__ void x3f_load(const GemmArgs& a, uint64_t r0, uint64_t c0, uint64_t k0,
                                         float (&fa)[X3F32Shape<BM, BN, WGM, WGN>::NPA][16],
                                         float (&fb)[X3F32Shape<BM, BN, WGM, WGN>::NPB][16])
{
    using F = X3F32Shape<BM, BN, WGM, WGN>;
    static_assert(F::RP % 64 == 0, "B's k half is wave-uniform");
    const int t = threadIdx.x;
    const uint32_t wave = __builtin_amdgcn_readfirstlane((uint32_t)t >> 6);
    const bool kfull = FULL || k0 + 32 <= a.K;         // block-uniform
    // A: 32 rows per wave and pass
    const uint32_t kh = (uint32_t)(k0 + (uint64_t)(t & 1) * 16), rl = (uint32_t)(t & 63) >> 1;
#pragma unroll
    for (int p = 0; p < F::NPA; p++) {
        const uint64_t wr = r0 + (uint64_t)wave * 32 + (uint64_t)p * F::RP;   // the wave's first row
        const uint64_t rows = wr < a.M ? (a.M - wr < 32 ? a.M - wr : 32) : 0;
        const __amdgpu_buffer_rsrc_t ra = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(a.Af + (wr < a.M ? wr : 0) * a.lda), (short)0,
            rows ? (int)(((rows - 1) * a.lda + a.K) * 4) : 0, 0x00020000);
        const uint32_t off = rl < rows ? (uint32_t)(rl * a.lda + kh) * 4 : 0x80000000u;
        if (AVEC && kfull) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const u32x4 x = __builtin_amdgcn_raw_buffer_load_b128(ra, off + 16 * q, 0, 0);
#pragma unroll
                for (int e = 0; e < 4; e++) fa[p][4 * q + e] = __uint_as_float(x[e]);
            }
        } else {
#pragma unroll
            for (int j = 0; j < 16; j++) {
                const uint32_t o = kh + j < a.K ? off + 4 * j : 0x80000000u;
                fa[p][j] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(ra, o, 0, 0));
            }
        }
    }
    // B: the stage's k rows (none for a stage entirely past K)
    if (!FULL && k0 >= a.K) {
#pragma unroll
        for (int p = 0; p < F::NPB; p++)
#pragma unroll
            for (int j = 0; j < 16; j++) fb[p][j] = 0.0f;
        return;
    }
    const uint32_t kbl = (uint32_t)(t / F::RP) * 16;   // wave-uniform
    const uint64_t krows = a.K - k0 < 32 ? a.K - k0 : 32;
    const __amdgpu_buffer_rsrc_t rb = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(a.Bf + k0 * a.ldb), (short)0, (int)(((krows - 1) * a.ldb + a.N) * 4), 0x00020000);
    if (c0 + (uint64_t)BN <= a.N) {
#pragma unroll
        for (int p = 0; p < F::NPB; p++) {
            const uint32_t off = (uint32_t)(kbl * a.ldb + c0 + (uint64_t)(t % F::RP) + (uint64_t)p * F::RP) * 4;
#pragma unroll
            for (int j = 0; j < 16; j++)
                fb[p][j] = kfull || kbl + j < krows
                               ? __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rb, off, (int)(j * a.ldb * 4), 0))
                               : 0.0f;
        }
    } else {
#pragma unroll
        for (int p = 0; p < F::NPB; p++) {
            const uint64_t col = c0 + (uint64_t)(t % F::RP) + (uint64_t)p * F::RP;
            const bool ok = col < a.N;
            const uint32_t off = (uint32_t)(kbl * a.ldb + (ok ? col : a.N - 1)) * 4;
#pragma unroll
            for (int j = 0; j < 16; j++) {
                float v = 0.0f;
                if (kfull || kbl + j < krows)
                    v = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rb, off, (int)(j * a.ldb * 4), 0));
                fb[p][j] = ok ? v : 0.0f;
            }
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

// x3f_tile for A rows that do (AVEC) or do not allow 16-byte loads.
template <int BM, int BN, int WGM, int WGN, bool PIPE, bool AVEC>
__device__ __forceinline__ void x3f_tile_body(const GemmArgs& a, uint32_t tm, uint32_t tn, u32x4* lds)
{
    using F = X3F32Shape<BM, BN, WGM, WGN>;
    using S = typename F::S;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = (wave / WGN) * S::WM, wn = (wave % WGN) * S::WN;
    const uint64_t r0 = (uint64_t)tm * BM, c0 = (uint64_t)tn * BN;

    floatx16 acc[S::TI][S::TJ];
#pragma unroll
    for (int i = 0; i < S::TI; i++)
#pragma unroll
        for (int j = 0; j < S::TJ; j++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[i][j][r] = 0.0f;

    float fa[F::NPA][16], fb[F::NPB][16];
    const int nk = (int)(a.Kp / 32);
    x3f_load<BM, BN, WGM, WGN, false, AVEC>(a, r0, c0, 0, fa, fb);
    x3f_store_stage<BM, BN, WGM, WGN>(lds, fa, fb);
    __syncthreads();

    int offA[S::TI][S::KB], offB[S::TJ][S::KB];
    x3_frag_offsets<S>(wm, wn, lane, offA, offB);

    if constexpr (!PIPE) {
        // Two f32 stages in registers: stage s sits in slot s & 1, is loaded
        // at step s - 3 and split into LDS at step s - 1, so each load has two
        // steps of MFMAs to arrive.  nk is even (Kp is a multiple of 64).
        float ga[F::NPA][16], gb[F::NPB][16];            // slot 1 (fa, fb: slot 0)
        x3f_load<BM, BN, WGM, WGN, false, AVEC>(a, r0, c0, 32, ga, gb);
        if (2 < nk) x3f_load<BM, BN, WGM, WGN, false, AVEC>(a, r0, c0, 64, fa, fb);
        // steps kt, kt + 1; the loads use x3f_load<FULL>
        auto two_steps = [&](int kt, auto full) {
            constexpr bool FULL = decltype(full)::value;
            // step kt: stage kt from buffer 0; stage kt + 1 (slot 1) into buffer 1
            x3_stage_mfma<S, 0>(lds, 0, offA, offB, acc);
            x3f_store_stage<BM, BN, WGM, WGN>(lds + S::STAGE, ga, gb);
            if (FULL || kt + 3 < nk) x3f_load<BM, BN, WGM, WGN, FULL, AVEC>(a, r0, c0, (uint64_t)(kt + 3) * 32, ga, gb);
            __syncthreads();
            // step kt + 1: stage kt + 1 from buffer 1; stage kt + 2 (slot 0) into buffer 0
            x3_stage_mfma<S, 0>(lds, S::STAGE, offA, offB, acc);
            if (FULL || kt + 2 < nk) {
                x3f_store_stage<BM, BN, WGM, WGN>(lds, fa, fb);
                if (FULL || kt + 4 < nk)
                    x3f_load<BM, BN, WGM, WGN, FULL, AVEC>(a, r0, c0, (uint64_t)(kt + 4) * 32, fa, fb);
            }
            __syncthreads();
        };
        // while both loads of a pair of steps are full stages, then the rest
        const int kfs = (int)(a.K / 32);
        int kt = 0;
        for (; kt + 4 < kfs; kt += 2) two_steps(kt, std::true_type{});
        for (; kt < nk; kt += 2) two_steps(kt, std::false_type{});
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
        if (more) x3f_load<BM, BN, WGM, WGN, false, AVEC>(a, r0, c0, (uint64_t)(kt + 1) * 32, fa, fb);
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

// One BM x BN output tile (tm, tn) from f32 A (a.Af, a.lda) and B (a.Bf,
// a.ldb); a.Kp = K rounded up to 64.  LDS is free on entry and on exit.
// PIPE: software-pipeline the MFMA operand fragments across the barrier.
template <int BM, int BN, int WGM, int WGN, bool PIPE>
__device__ __forceinline__ void x3f_tile(const GemmArgs& a, uint32_t tm, uint32_t tn, u32x4* lds)
{
    if (((a.lda & 3) == 0) && (((uintptr_t)a.Af & 15) == 0))
        x3f_tile_body<BM, BN, WGM, WGN, PIPE, true>(a, tm, tn, lds);
    else
        x3f_tile_body<BM, BN, WGM, WGN, PIPE, false>(a, tm, tn, lds);
}

// The f32-fed family's tile shapes: X(BM, BN, waves_m, waves_n, min blocks per CU)
#define CHAI_X3F_VARIANTS(X)                                                  \
    X(128, 128, 2, 2, 2)   /* 4 waves of 64x64, 2 blocks / CU */             \
    X(256, 128, 2, 2, 1)   /* 4 waves of 128x64                */

}  // namespace chai_hip
