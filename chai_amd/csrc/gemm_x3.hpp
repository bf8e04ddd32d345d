// gemm_x3.hpp -- split-BF16x3 MFMA GEMM core, templated over its tile shape.
//
// C = A B with A = Ah + Al and B = Bh + Bl (bf16 pairs from the split
// pre-pass; A as Mp x Kp and B^T as Np x Kp panels, both in k blocks of 32:
// element (r, k) at (k / 32) * Rp * 32 + r * 32 + k % 32).
// Each 32x32 output tile accumulates, per 16-deep k block, the three products
// Ah*Bl + Al*Bh + Ah*Bh with v_mfma_f32_32x32x16_bf16 into one FP32
// accumulator (the small terms first).
//
// Tile parameters (the autotuner's search space):
//   BM x BN   block tile, BK k depth per LDS stage (32 or 64 bf16)
//   WGM x WGN waves per block; each wave owns a (BM/WGM) x (BN/WGN) tile
// LDS holds two stages of the four panels Ah, Al, Bh, Bl; rows are 16-byte
// chunks XOR-swizzled so that the rows one MFMA operand read touches fall
// into distinct banks.  Stages are copied from global memory straight into
// LDS (global_load_lds_dwordx4: lane i of a wave fills the i-th 16 bytes
// after the wave's LDS base, so each lane fetches the source chunk that the
// swizzle puts at its position): stage k+2 is issued right after the barrier
// that frees its buffer and lands while stage k+1 is multiplied; one vmcnt
// wait and one barrier per stage.  No staging registers, no ds_write.
// The kernel body avoids lambdas and keeps every per-thread array fully
// unrolled so that nothing lands in scratch.
#pragma once
#include "gemm_device.hpp"
#include "gemm_internal.hpp"

namespace chai_hip {

__device__ __forceinline__ bf16x8 as_bf16x8(u32x4 v)
{
    bf16x8 r;
    __builtin_memcpy(&r, &v, 16);
    return r;
}

__device__ __forceinline__ void store_c(float* C, uint64_t idx, float v, const GemmArgs& a)
{
    if (a.general) v = (float)(a.alpha * (double)v + a.beta * (double)gload(C + idx));
    if (a.epi) v = fmaxf(v * a.epi_scale + a.epi_shift, a.epi_lo);   // no contraction (-ffp-contract=off)
    gstore(C + idx, v);
}

template <int BM, int BN, int BK, int WGM, int WGN>
struct X3Shape {
    static constexpr int THREADS = WGM * WGN * 64;
    static constexpr int WM = BM / WGM, WN = BN / WGN;
    static constexpr int TI = WM / 32, TJ = WN / 32;
    static constexpr int CPR = BK / 8;                 // 16-byte chunks per panel row
    static constexpr int SH = BK == 32 ? 2 : 1;        // rows sharing one 256-byte bank line
    static constexpr int PA = BM * CPR, PB = BN * CPR; // chunks per panel per stage
    static constexpr int STAGE = 2 * PA + 2 * PB;
    static constexpr int WAVES = THREADS / 64;
    static constexpr int IA = PA / 64 / WAVES, IB = PB / 64 / WAVES;   // LDS-DMA issues per panel
    static constexpr int KB = BK / 16;
    static constexpr size_t LDS_BYTES = 2 * (size_t)STAGE * 16;
    // LDS reads and MFMAs of one k block, without the zero panels (SKIP)
    template <int SKIP> static constexpr int frag_reads() { return TI * ((SKIP & 1) ? 1 : 2) + TJ * ((SKIP & 2) ? 1 : 2); }
    template <int SKIP> static constexpr int frag_mfmas() { return TI * TJ * (1 + !(SKIP & 1) + !(SKIP & 2)); }
    static_assert(WM % 32 == 0 && WN % 32 == 0, "wave tile must be a multiple of 32");
    static_assert(BK == 32 || BK == 64, "BK is 32 or 64");
    static_assert(PA % THREADS == 0 && PB % THREADS == 0, "panel chunks must split evenly");
    static_assert(IA * 64 * WAVES == PA && IB * 64 * WAVES == PB, "whole wave issues per panel");
};

// Issue the LDS-DMA copies of one stage into the stage buffer st; wave w
// copies chunks [(w * I + s) * 64, +64) of each panel.  The panel pointers
// are wave-uniform (tile origin and stage k offset folded in); a lane adds its
// 32-bit offset within the k block and, for a 64-deep stage, one k block when
// bit s of its mask is set.
template <typename S, int SKIP>
__device__ __forceinline__ void x3_issue_stage(const uint16_t* pAh, const uint16_t* pAl, const uint16_t* pBh,
                                               const uint16_t* pBl, const uint32_t (&srcA)[S::IA],
                                               const uint32_t (&srcB)[S::IB], uint32_t hiA, uint32_t hiB,
                                               uint64_t a_kblk, uint64_t b_kblk, u32x4* st, int wave)
{
#pragma unroll
    for (int s = 0; s < S::IA; s++) {
        const int c0 = (wave * S::IA + s) * 64;
        const uint64_t o = srcA[s] + (S::CPR == 8 && ((hiA >> s) & 1u) ? a_kblk : 0);
        // panel base (wave-uniform) + the lane's 32-bit byte offset
        const uint32_t ob = (uint32_t)(o * 2);
        lds_dma16(reinterpret_cast<const char*>(pAh) + ob, st + c0);
        if (!(SKIP & 1)) lds_dma16(reinterpret_cast<const char*>(pAl) + ob, st + S::PA + c0);
    }
#pragma unroll
    for (int s = 0; s < S::IB; s++) {
        const int c0 = (wave * S::IB + s) * 64;
        const uint64_t o = srcB[s] + (S::CPR == 8 && ((hiB >> s) & 1u) ? b_kblk : 0);
        const uint32_t ob = (uint32_t)(o * 2);
        lds_dma16(reinterpret_cast<const char*>(pBh) + ob, st + 2 * S::PA + c0);
        if (!(SKIP & 2)) lds_dma16(reinterpret_cast<const char*>(pBl) + ob, st + 2 * S::PA + S::PB + c0);
    }
}

// Fragment read offsets (u32x4 units within a stage) of this lane's A and B
// tile rows for each 16-deep k block.
template <typename S>
__device__ __forceinline__ void x3_frag_offsets(int wm, int wn, int lane, int (&offA)[S::TI][S::KB],
                                                int (&offB)[S::TJ][S::KB])
{
#pragma unroll
    for (int kb = 0; kb < S::KB; kb++) {
        const int q = kb * 2 + (lane >> 5);
#pragma unroll
        for (int i = 0; i < S::TI; i++) {
            const int row = wm + i * 32 + (lane & 31);
            offA[i][kb] = row * S::CPR + (q ^ ((row >> S::SH) & (S::CPR - 1)));
        }
#pragma unroll
        for (int j = 0; j < S::TJ; j++) {
            const int row = wn + j * 32 + (lane & 31);
            offB[j][kb] = 2 * S::PA + row * S::CPR + (q ^ ((row >> S::SH) & (S::CPR - 1)));
        }
    }
}

// One 16-deep k block's operand fragments of a wave: hi and lo of its TI A
// tiles and TJ B tiles (a register slot of the fragment pipeline).
template <typename S>
struct X3Frag {
    bf16x8 ah[S::TI], al[S::TI], bh[S::TJ], bl[S::TJ];
};

// Read k block kb of the LDS stage at chunk offset cur into f, minus the zero
// panels (SKIP).
template <typename S, int SKIP>
__device__ __forceinline__ void x3_load_frag(const u32x4* lds, int cur, int kb, const int (&offA)[S::TI][S::KB],
                                             const int (&offB)[S::TJ][S::KB], X3Frag<S>& f)
{
#pragma unroll
    for (int i = 0; i < S::TI; i++) {
        f.ah[i] = as_bf16x8(lds[cur + offA[i][kb]]);
        if (!(SKIP & 1)) f.al[i] = as_bf16x8(lds[cur + S::PA + offA[i][kb]]);
    }
#pragma unroll
    for (int j = 0; j < S::TJ; j++) {
        f.bh[j] = as_bf16x8(lds[cur + offB[j][kb]]);
        if (!(SKIP & 2)) f.bl[j] = as_bf16x8(lds[cur + S::PB + offB[j][kb]]);
    }
}

// The MFMAs of one k block: the three products ah*bl, al*bh, ah*bh (minus
// those with a zero panel), each as a pass over the tiles so that consecutive
// MFMAs write different accumulators.
template <typename S, int SKIP>
__device__ __forceinline__ void x3_frag_mfma(const X3Frag<S>& f, floatx16 (&acc)[S::TI][S::TJ])
{
    if (!(SKIP & 2)) {
#pragma unroll
        for (int i = 0; i < S::TI; i++)
#pragma unroll
            for (int j = 0; j < S::TJ; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.ah[i], f.bl[j], acc[i][j], 0, 0, 0);
    }
    if (!(SKIP & 1)) {
#pragma unroll
        for (int i = 0; i < S::TI; i++)
#pragma unroll
            for (int j = 0; j < S::TJ; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.al[i], f.bh[j], acc[i][j], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < S::TI; i++)
#pragma unroll
        for (int j = 0; j < S::TJ; j++)
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f.ah[i], f.bh[j], acc[i][j], 0, 0, 0);
}

// All MFMAs of one LDS stage (at chunk offset cur), k block by k block.
template <typename S, int SKIP>
__device__ __forceinline__ void x3_stage_mfma(const u32x4* lds, int cur, const int (&offA)[S::TI][S::KB],
                                              const int (&offB)[S::TJ][S::KB], floatx16 (&acc)[S::TI][S::TJ])
{
#pragma unroll
    for (int kb = 0; kb < S::KB; kb++) {
        X3Frag<S> f;
        x3_load_frag<S, SKIP>(lds, cur, kb, offA, offB, f);
        x3_frag_mfma<S, SKIP>(f, acc);
    }
}

// Store the wave's accumulators (32x32x16 output layout: lane -> column
// lane % 32; register r -> row 8 (r / 4) + 4 (lane / 32) + r % 4).  The
// general (alpha / beta) and epilogue forms are separate loops, so each
// fully unrolled loop stays small enough to keep the accumulators in
// registers.
template <typename S, bool GEN, bool EPI>
__device__ __forceinline__ void x3_store_as(const GemmArgs& a, float* C, uint64_t row_base, uint64_t col_base,
                                            int lane, const floatx16 (&acc)[S::TI][S::TJ])
{
    const bool full = row_base + S::TI * 32 <= a.M && col_base + S::TJ * 32 <= a.N;
#pragma unroll
    for (int i = 0; i < S::TI; i++)
#pragma unroll
        for (int j = 0; j < S::TJ; j++) {
            const uint64_t col = col_base + j * 32 + (lane & 31);
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const uint64_t row = row_base + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                if (full || (row < a.M && col < a.N)) {
                    float v = acc[i][j][r];
                    float* p = C + row * a.ldc + col;
                    if (GEN) v = (float)(a.alpha * (double)v + a.beta * (double)gload(p));
                    if (EPI) v = fmaxf(v * a.epi_scale + a.epi_shift, a.epi_lo);   // no contraction (-ffp-contract=off)
                    gstore(p, v);
                }
            }
        }
}

template <typename S>
__device__ __forceinline__ void x3_store(const GemmArgs& a, float* C, uint64_t row_base, uint64_t col_base, int lane,
                                         const floatx16 (&acc)[S::TI][S::TJ])
{
    if (a.general) {
        if (a.epi) x3_store_as<S, true, true>(a, C, row_base, col_base, lane, acc);
        else x3_store_as<S, true, false>(a, C, row_base, col_base, lane, acc);
    } else {
        if (a.epi) x3_store_as<S, false, true>(a, C, row_base, col_base, lane, acc);
        else x3_store_as<S, false, false>(a, C, row_base, col_base, lane, acc);
    }
}

// The tile body.  SKIP bit 0: the tile's Al panel is zero, bit 1: its Bl
// panel is zero; the products with a zero panel (exact zeros) are neither
// loaded nor multiplied.
template <int BM, int BN, int BK, int WGM, int WGN, int SKIP>
__device__ __forceinline__ void x3_tile_body(const GemmArgs& a, uint32_t tm, uint32_t tn, uint64_t bz, u32x4* lds)
{
    using S = X3Shape<BM, BN, BK, WGM, WGN>;
    const uint64_t Kp = a.Kp;
    float* C = a.C + bz * a.c_bstride;

    // the wave index in a scalar register: the LDS destinations of the
    // stage copies (M0) and the panel offsets derived from it stay scalar
    const int t = threadIdx.x, lane = t & 63, wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const int wm = (wave / WGN) * S::WM;
    const int wn = (wave % WGN) * S::WN;

    // LDS chunk c = (wave * IA + s) * 64 + lane of a panel is row c / CPR,
    // position c % CPR, and holds source chunk q = position ^ swizzle(row):
    // k block q / 4 of the stage, 8 elements at (q % 4) * 8 within it.
    uint32_t srcA[S::IA], srcB[S::IB];   // offsets within the k block, from the tile origin
    uint32_t hiA = 0, hiB = 0;            // bit s: chunk s lies in the stage's second k block
#pragma unroll
    for (int s = 0; s < S::IA; s++) {
        const int c = (wave * S::IA + s) * 64 + lane, row = c / S::CPR, p = c % S::CPR;
        const int q = p ^ ((row >> S::SH) & (S::CPR - 1));
        srcA[s] = (uint32_t)(row * 32 + (q & 3) * 8);
        hiA |= (uint32_t)(q >> 2) << s;
    }
#pragma unroll
    for (int s = 0; s < S::IB; s++) {
        const int c = (wave * S::IB + s) * 64 + lane, row = c / S::CPR, p = c % S::CPR;
        const int q = p ^ ((row >> S::SH) & (S::CPR - 1));
        srcB[s] = (uint32_t)(row * 32 + (q & 3) * 8);
        hiB |= (uint32_t)(q >> 2) << s;
    }
    const uint64_t a0 = bz * a.pa_bstride + (uint64_t)tm * BM * 32;   // tile origin, k block 0
    const uint64_t b0 = bz * a.pb_bstride + (uint64_t)tn * BN * 32;
    const uint64_t kstepA = (uint64_t)(BK / 32) * a.a_kblk, kstepB = (uint64_t)(BK / 32) * a.b_kblk;

    floatx16 acc[S::TI][S::TJ];
#pragma unroll
    for (int i = 0; i < S::TI; i++)
#pragma unroll
        for (int j = 0; j < S::TJ; j++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[i][j][r] = 0.0f;

    const int nk = (int)(Kp / BK);
    // prologue: stage 0 -> LDS buffer 0, visible to all waves; stage 1 -> buffer 1
    x3_issue_stage<S, SKIP>(a.Ah + a0, a.Al + a0, a.Bh + b0, a.Bl + b0, srcA, srcB, hiA, hiB, a.a_kblk,
                            a.b_kblk, lds, wave);
    wait_lds_dma();
    __syncthreads();
    {
        const uint64_t k1 = nk > 1 ? 1 : 0;
        x3_issue_stage<S, SKIP>(a.Ah + a0 + k1 * kstepA, a.Al + a0 + k1 * kstepA, a.Bh + b0 + k1 * kstepB,
                                a.Bl + b0 + k1 * kstepB, srcA, srcB, hiA, hiB, a.a_kblk, a.b_kblk,
                                lds + S::STAGE, wave);
    }

    int offA[S::TI][S::KB], offB[S::TJ][S::KB];
    x3_frag_offsets<S>(wm, wn, lane, offA, offB);

    // Fragment pipeline: the registers of k block kb + 1 are read while the
    // MFMAs of k block kb run.  The last k block of stage kt is multiplied
    // after the stage barrier, beside the reads of stage kt + 1's first k
    // block and the copy of stage kt + 2, so the matrix pipe is fed across
    // the barrier.  KB is even, so k block kb always sits in slot kb & 1.
    // Scheduling fences keep each group of MFMAs on its side of the barrier.
    // The copy is issued every step (stage min(kt + 2, nk - 1)): in the last
    // two steps it refills a buffer no wave reads again, which keeps the
    // step free of branches.
    X3Frag<S> f[2];
    x3_load_frag<S, SKIP>(lds, 0, 0, offA, offB, f[0]);
    for (int kt = 0; kt < nk; kt++) {
        const int cur = (kt & 1) * S::STAGE;
        // k block 0 has arrived (its reads went out behind the last step's
        // MFMAs); waiting here leaves the reads below free to stay in flight
        __builtin_amdgcn_s_waitcnt(0xc07f);   // lgkmcnt(0)
#pragma unroll
        for (int kb = 0; kb + 1 < S::KB; kb++) {
            x3_load_frag<S, SKIP>(lds, cur, kb + 1, offA, offB, f[(kb + 1) & 1]);
            x3_frag_mfma<S, SKIP>(f[kb & 1], acc);
            // all reads of k block kb + 1 go out before the MFMAs of kb
            __builtin_amdgcn_sched_group_barrier(0x100, S::template frag_reads<SKIP>(), 0);
            __builtin_amdgcn_sched_group_barrier(0x008, S::template frag_mfmas<SKIP>(), 0);
        }
        __builtin_amdgcn_sched_barrier(0);
        // every wave's reads of buffer cur have completed (the barrier waits
        // for them) and stage kt + 1 has landed in the other buffer
        wait_lds_dma();
        __syncthreads();
        __builtin_amdgcn_sched_barrier(0);
        const int kn = kt + 1 < nk ? kt + 1 : kt;
        x3_load_frag<S, SKIP>(lds, (kn & 1) * S::STAGE, 0, offA, offB, f[0]);
        const uint64_t k2 = (uint64_t)(kt + 2 < nk ? kt + 2 : nk - 1);
        const uint64_t oa = a0 + k2 * kstepA, ob = b0 + k2 * kstepB;
        x3_issue_stage<S, SKIP>(a.Ah + oa, a.Al + oa, a.Bh + ob, a.Bl + ob, srcA, srcB, hiA, hiB, a.a_kblk,
                                a.b_kblk, lds + cur, wave);
        x3_frag_mfma<S, SKIP>(f[(S::KB - 1) & 1], acc);
    }
    wait_lds_dma();     // the refills of the last two steps have landed: LDS is free on exit

    x3_store<S>(a, C, (uint64_t)tm * BM + wm, (uint64_t)tn * BN + wn, lane, acc);
}

// One BM x BN output tile (tm, tn) of batch instance bz.  The block's LDS is
// free on entry and on exit (the last k step ends with a barrier), so a
// persistent kernel may call it for tile after tile.  With split flags the
// block first ORs the flags of its A and B panels (a block-uniform decision).
template <int BM, int BN, int BK, int WGM, int WGN>
__device__ __forceinline__ void x3_tile(const GemmArgs& a, uint32_t tm, uint32_t tn, uint64_t bz, u32x4* lds)
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
        case 0: x3_tile_body<BM, BN, BK, WGM, WGN, 0>(a, tm, tn, bz, lds); break;
        case 1: x3_tile_body<BM, BN, BK, WGM, WGN, 1>(a, tm, tn, bz, lds); break;
        case 2: x3_tile_body<BM, BN, BK, WGM, WGN, 2>(a, tm, tn, bz, lds); break;
        default: x3_tile_body<BM, BN, BK, WGM, WGN, 3>(a, tm, tn, bz, lds); break;
    }
}

// The autotuner's search space for the split-BF16x3 family, in variant-id
// order: X(BM, BN, BK, waves_m, waves_n, min blocks per CU).
#define CHAI_X3_VARIANTS(X)                                                   \
    X(256, 256, 32, 2, 4, 1)   /* 0: 8 waves of 128x64               */      \
    X(256, 256, 32, 4, 2, 1)   /* 1: 8 waves of 64x128               */      \
    X(128, 128, 32, 2, 2, 2)   /* 2: 4 waves of 64x64, 2 blocks / CU */      \
    X(256, 128, 32, 2, 2, 1)   /* 3: 4 waves of 128x64               */      \
    X(128, 256, 32, 2, 2, 1)   /* 4: 4 waves of 64x128               */      \
    X(128, 128, 64, 2, 2, 1)   /* 5: BK 64                           */      \
    X(128, 256, 32, 2, 4, 1)   /* 6: 8 waves of 64x64                */      \
    X(256, 128, 32, 4, 2, 1)   /* 7: 8 waves of 64x64                */

template <int BM, int BN, int BK, int WGM, int WGN, int MINB>
__global__ __launch_bounds__(WGM * WGN * 64, MINB) void gemm_x3_kernel(GemmArgs a)
{
    extern __shared__ u32x4 lds[];
    uint32_t tm, tn;
    tile_coords<8>(blockIdx.x, a.tiles_m, a.tiles_n, tm, tn);
    x3_tile<BM, BN, BK, WGM, WGN>(a, tm, tn, blockIdx.z, lds);
}

}  // namespace chai_hip
