// gemm_x6.hpp -- split-BF16x6 MFMA GEMM core: three bf16 pieces per operand,
// six products, one pass.
//
// A = a0 + a1 + a2 and B = b0 + b1 + b2 (bf16 pieces from the split3
// pre-pass: a0 = bf16(a), a1 = bf16(a - a0), a2 = bf16(a - a0 - a1), so the
// three carry all 24 bits of an f32).  Each 32x32 output tile accumulates, per
// 16-deep k block, the six products of order <= 2 with
// v_mfma_f32_32x32x16_bf16 into one FP32 accumulator, small terms first:
//   a1 b1, a0 b2, a2 b0   (order 2, below 2^-16 |a b|)
//   a0 b1, a1 b0          (order 1)
//   a0 b0                 (order 0)
// The terms dropped (a1 b2, a2 b1, a2 b2) are below 2^-24 |a b|, so the error
// is FP32-FMA level at twice the BF16X3 MFMA work.
//
// Panels: A pieces as Mp x Kp and B^T pieces as Np x Kp panels in k blocks of
// 32 (gemm_x3.hpp's layout); the three pieces of an operand are consecutive
// panels (piece p at p * Rp * Kp).  LDS holds two stages of the six panels,
// rows as 16-byte chunks XOR-swizzled, filled straight from global memory
// (global_load_lds_dwordx4); the operand fragments are software-pipelined
// across the stage barrier as in the BF16X3 tile.  With the split flags a
// tile whose a1 (b1) panel is all zero -- then so is a2 (b2) -- skips the
// products with either.
#pragma once
#include "gemm_x3.hpp"

namespace chai_hip {

template <int BM, int BN, int BK, int WGM, int WGN>
struct X6Shape {
    static constexpr int THREADS = WGM * WGN * 64;
    static constexpr int WM = BM / WGM, WN = BN / WGN;
    static constexpr int TI = WM / 32, TJ = WN / 32;
    static constexpr int CPR = BK / 8;                 // 16-byte chunks per panel row
    static constexpr int SH = BK == 32 ? 2 : 1;
    static constexpr int PA = BM * CPR, PB = BN * CPR; // chunks per panel per stage
    static constexpr int STAGE = 3 * PA + 3 * PB;
    static constexpr int WAVES = THREADS / 64;
    static constexpr int IA = PA / 64 / WAVES, IB = PB / 64 / WAVES;
    static constexpr int KB = BK / 16;
    static constexpr size_t LDS_BYTES = 2 * (size_t)STAGE * 16;
    // LDS reads and MFMAs of one k block, without the zero panels (SKIP)
    template <int SKIP> static constexpr int frag_reads() { return TI * ((SKIP & 1) ? 1 : 3) + TJ * ((SKIP & 2) ? 1 : 3); }
    template <int SKIP> static constexpr int frag_mfmas()
    {
        return TI * TJ * (SKIP == 0 ? 6 : SKIP == 3 ? 1 : 3);
    }
    static_assert(WM % 32 == 0 && WN % 32 == 0, "wave tile must be a multiple of 32");
    static_assert(BK == 32, "32-deep stages");
    static_assert(PA % THREADS == 0 && PB % THREADS == 0, "panel chunks must split evenly");
    static_assert(IA * 64 * WAVES == PA && IB * 64 * WAVES == PB, "whole wave issues per panel");
    static_assert(LDS_BYTES <= 160 * 1024, "two stages fit the LDS");
};

// LDS-DMA copies of one stage: wave w copies chunks [(w * I + s) * 64, +64)
// of each panel; p*[q] is piece q's panel at the stage's k offset.
template <typename S, int SKIP>
__device__ __forceinline__ void x6_issue_stage(const uint16_t* const (&pa)[3], const uint16_t* const (&pb)[3],
                                               const uint32_t (&srcA)[S::IA], const uint32_t (&srcB)[S::IB],
                                               u32x4* st, int wave)
{
#pragma unroll
    for (int s = 0; s < S::IA; s++) {
        const int c0 = (wave * S::IA + s) * 64;
#pragma unroll
        for (int q = 0; q < 3; q++)
            if (q == 0 || !(SKIP & 1)) lds_dma16(pa[q] + srcA[s], st + q * S::PA + c0);
    }
#pragma unroll
    for (int s = 0; s < S::IB; s++) {
        const int c0 = (wave * S::IB + s) * 64;
#pragma unroll
        for (int q = 0; q < 3; q++)
            if (q == 0 || !(SKIP & 2)) lds_dma16(pb[q] + srcB[s], st + 3 * S::PA + q * S::PB + c0);
    }
}

template <typename S>
struct X6Frag {
    bf16x8 a[3][S::TI], b[3][S::TJ];
};

template <typename S, int SKIP>
__device__ __forceinline__ void x6_load_frag(const u32x4* lds, int cur, int kb, const int (&offA)[S::TI][S::KB],
                                             const int (&offB)[S::TJ][S::KB], X6Frag<S>& f)
{
#pragma unroll
    for (int q = 0; q < 3; q++)
        if (q == 0 || !(SKIP & 1))
#pragma unroll
            for (int i = 0; i < S::TI; i++) f.a[q][i] = as_bf16x8(lds[cur + q * S::PA + offA[i][kb]]);
#pragma unroll
    for (int q = 0; q < 3; q++)
        if (q == 0 || !(SKIP & 2))
#pragma unroll
            for (int j = 0; j < S::TJ; j++) f.b[q][j] = as_bf16x8(lds[cur + q * S::PB + offB[j][kb]]);
}

// acc += a_p b_q over the wave's tiles (one pass: consecutive MFMAs write
// different accumulators)
template <typename S, int P, int Q>
__device__ __forceinline__ void x6_pass(const X6Frag<S>& f, floatx16 (&acc)[S::TI][S::TJ])
{
#pragma unroll
    for (int i = 0; i < S::TI; i++)
#pragma unroll
        for (int j = 0; j < S::TJ; j++)
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.a[P][i], f.b[Q][j], acc[i][j], 0, 0, 0);
}

template <typename S, int SKIP>
__device__ __forceinline__ void x6_frag_mfma(const X6Frag<S>& f, floatx16 (&acc)[S::TI][S::TJ])
{
    if (SKIP == 0) {
        x6_pass<S, 1, 1>(f, acc);
        x6_pass<S, 0, 2>(f, acc);
        x6_pass<S, 2, 0>(f, acc);
    }
    if (!(SKIP & 1)) x6_pass<S, 1, 0>(f, acc);    // a1 b0 (order 1) ...
    if (!(SKIP & 2)) x6_pass<S, 0, 1>(f, acc);    // ... a0 b1
    if (SKIP == 1) x6_pass<S, 0, 2>(f, acc);      // a lo pieces zero: a0 b2 is the order-2 term left
    if (SKIP == 2) x6_pass<S, 2, 0>(f, acc);      // b lo pieces zero: a2 b0
    x6_pass<S, 0, 0>(f, acc);
}

template <int BM, int BN, int BK, int WGM, int WGN, int SKIP>
__device__ __forceinline__ void x6_tile_body(const GemmArgs& a, uint32_t tm, uint32_t tn, uint64_t bz, u32x4* lds)
{
    using S = X6Shape<BM, BN, BK, WGM, WGN>;
    using S3 = X3Shape<BM, BN, BK, WGM, WGN>;     // same fragment geometry and store layout
    float* C = a.C + bz * a.c_bstride;
    // wave index in a scalar register: the copies' LDS destinations stay scalar
    const int t = threadIdx.x, lane = t & 63, wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const int wm = (wave / WGN) * S::WM;
    const int wn = (wave % WGN) * S::WN;

    // LDS chunk c = (wave * IA + s) * 64 + lane of a panel is row c / CPR,
    // position c % CPR, and holds source chunk (position ^ swizzle(row)) of
    // the row's 32 k
    uint32_t srcA[S::IA], srcB[S::IB];
#pragma unroll
    for (int s = 0; s < S::IA; s++) {
        const int c = (wave * S::IA + s) * 64 + lane, row = c / S::CPR, p = c % S::CPR;
        srcA[s] = (uint32_t)(row * 32 + (p ^ ((row >> S::SH) & (S::CPR - 1))) * 8);
    }
#pragma unroll
    for (int s = 0; s < S::IB; s++) {
        const int c = (wave * S::IB + s) * 64 + lane, row = c / S::CPR, p = c % S::CPR;
        srcB[s] = (uint32_t)(row * 32 + (p ^ ((row >> S::SH) & (S::CPR - 1))) * 8);
    }
    const uint64_t a0 = bz * a.pa_bstride + (uint64_t)tm * BM * 32;   // tile origin, k block 0
    const uint64_t b0 = bz * a.pb_bstride + (uint64_t)tn * BN * 32;
    const uint64_t pieceA = a.Mp * a.Kp, pieceB = a.Np * a.Kp;          // elements between pieces
    const uint64_t kstepA = a.a_kblk, kstepB = a.b_kblk;                // one 32-deep stage

    floatx16 acc[S::TI][S::TJ];
#pragma unroll
    for (int i = 0; i < S::TI; i++)
#pragma unroll
        for (int j = 0; j < S::TJ; j++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[i][j][r] = 0.0f;

    const int nk = (int)(a.Kp / BK);
    auto issue = [&](uint64_t k, u32x4* st) {
        const uint64_t oa = a0 + k * kstepA, ob = b0 + k * kstepB;
        const uint16_t* const pa[3] = {a.Ah + oa, a.Ah + pieceA + oa, a.Ah + 2 * pieceA + oa};
        const uint16_t* const pb[3] = {a.Bh + ob, a.Bh + pieceB + ob, a.Bh + 2 * pieceB + ob};
        x6_issue_stage<S, SKIP>(pa, pb, srcA, srcB, st, wave);
    };
    issue(0, lds);
    wait_lds_dma();
    __syncthreads();
    issue(nk > 1 ? 1 : 0, lds + S::STAGE);

    int offA[S::TI][S::KB], offB[S::TJ][S::KB];
    x3_frag_offsets<S3>(wm, wn, lane, offA, offB);
    // x3_frag_offsets places B after two A panels; here B follows three
#pragma unroll
    for (int j = 0; j < S::TJ; j++)
#pragma unroll
        for (int kb = 0; kb < S::KB; kb++) offB[j][kb] += S::PA;

    X6Frag<S> f[2];
    x6_load_frag<S, SKIP>(lds, 0, 0, offA, offB, f[0]);
    for (int kt = 0; kt < nk; kt++) {
        const int cur = (kt & 1) * S::STAGE;
        __builtin_amdgcn_s_waitcnt(0xc07f);   // lgkmcnt(0): k block 0 has arrived
        x6_load_frag<S, SKIP>(lds, cur, 1, offA, offB, f[1]);
        x6_frag_mfma<S, SKIP>(f[0], acc);
        __builtin_amdgcn_sched_group_barrier(0x100, S::template frag_reads<SKIP>(), 0);
        __builtin_amdgcn_sched_group_barrier(0x008, S::template frag_mfmas<SKIP>(), 0);
        __builtin_amdgcn_sched_barrier(0);
        wait_lds_dma();
        __syncthreads();
        __builtin_amdgcn_sched_barrier(0);
        const int kn = kt + 1 < nk ? kt + 1 : kt;
        x6_load_frag<S, SKIP>(lds, (kn & 1) * S::STAGE, 0, offA, offB, f[0]);
        issue((uint64_t)(kt + 2 < nk ? kt + 2 : nk - 1), lds + cur);
        x6_frag_mfma<S, SKIP>(f[1], acc);
    }
    wait_lds_dma();     // the refills of the last two steps have landed: LDS is free on exit

    x3_store<S3>(a, C, (uint64_t)tm * BM + wm, (uint64_t)tn * BN + wn, lane, acc);
}

template <int BM, int BN, int BK, int WGM, int WGN>
__device__ __forceinline__ void x6_tile(const GemmArgs& a, uint32_t tm, uint32_t tn, uint64_t bz, u32x4* lds)
{
    int skip = 0;
    if (a.a_nz) {
        constexpr int THREADS = WGM * WGN * 64;
        const uint64_t kb64 = a.Kp / 64;
        const uint64_t na = (uint64_t)(BM / 64) * kb64, nb = (uint64_t)(BN / 64) * kb64;
        const uint8_t* fa = a.a_nz + bz * a.a_nz_bstride + (uint64_t)tm * na;
        const uint8_t* fb = a.b_nz + bz * a.b_nz_bstride + (uint64_t)tn * nb;
        int nz_a = 0, nz_b = 0;
        for (uint64_t i = threadIdx.x; i < na; i += THREADS) nz_a |= fa[i];
        for (uint64_t i = threadIdx.x; i < nb; i += THREADS) nz_b |= fb[i];
        skip = (__syncthreads_or(nz_a) ? 0 : 1) | (__syncthreads_or(nz_b) ? 0 : 2);
    }
    switch (skip) {
        case 0: x6_tile_body<BM, BN, BK, WGM, WGN, 0>(a, tm, tn, bz, lds); break;
        case 1: x6_tile_body<BM, BN, BK, WGM, WGN, 1>(a, tm, tn, bz, lds); break;
        case 2: x6_tile_body<BM, BN, BK, WGM, WGN, 2>(a, tm, tn, bz, lds); break;
        default: x6_tile_body<BM, BN, BK, WGM, WGN, 3>(a, tm, tn, bz, lds); break;
    }
}

// The autotuner's search space for the split-BF16x6 family, in variant-id
// order: X(BM, BN, BK, waves_m, waves_n, min blocks per CU).  Two stages of
// six panels must fit the 160 KB LDS.
#define CHAI_X6_VARIANTS(X)                                                   \
    X(256, 128, 32, 4, 2, 1)   /* 0: 8 waves of 64x64                */      \
    X(128, 256, 32, 2, 4, 1)   /* 1: 8 waves of 64x64                */      \
    X(128, 128, 32, 2, 2, 1)   /* 2: 4 waves of 64x64                */      \
    X(256, 128, 32, 2, 2, 1)   /* 3: 4 waves of 128x64               */      \
    X(128, 256, 32, 2, 2, 1)   /* 4: 4 waves of 64x128               */

template <int BM, int BN, int BK, int WGM, int WGN, int MINB>
__global__ __launch_bounds__(WGM * WGN * 64, MINB) void gemm_x6_kernel(GemmArgs a)
{
    extern __shared__ u32x4 lds[];
    uint32_t tm, tn;
    tile_coords<8>(blockIdx.x, a.tiles_m, a.tiles_n, tm, tn);
    x6_tile<BM, BN, BK, WGM, WGN>(a, tm, tn, blockIdx.z, lds);
}

}  // namespace chai_hip
