// gemm_x3.hpp -- split-BF16x3 MFMA GEMM core, templated over its tile shape.
//
// C = A B with A = Ah + Al and B = Bh + Bl (bf16 pairs from the split
// pre-pass, both stored K-contiguous: A as [Mp][Kp], B^T as [Np][Kp]).
// Each 32x32 output tile accumulates, per 16-deep k block, the three products
// Ah*Bl + Al*Bh + Ah*Bh with v_mfma_f32_32x32x16_bf16 into one FP32
// accumulator (the small terms first).
//
// Tile parameters (the autotuner's search space):
//   BM x BN   block tile, BK k depth per LDS stage (32 or 64 bf16)
//   WGM x WGN waves per block; each wave owns a (BM/WGM) x (BN/WGN) tile
// LDS holds two stages of the four panels Ah, Al, Bh, Bl; rows are 16-byte
// chunks XOR-swizzled so that the rows one MFMA operand read touches fall
// into distinct banks.  Global loads for stage k+1 are issued into registers
// before the MFMAs of stage k and written to LDS after them (one barrier per
// stage); fragments of the next 16-deep k block are read from LDS while the
// MFMAs of the current one run.  The kernel body avoids lambdas and keeps
// every per-thread array fully unrolled so that nothing lands in scratch.
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
    if (a.general) v = (float)(a.alpha * (double)v + a.beta * (double)C[idx]);
    if (a.epi) v = fmaxf(v * a.epi_scale + a.epi_shift, a.epi_lo);   // no contraction (-ffp-contract=off)
    C[idx] = v;
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
    static constexpr int NA = PA / THREADS, NB = PB / THREADS;   // load slots per panel
    static constexpr int RA = THREADS / CPR;           // rows covered by one load slot
    static constexpr int KB = BK / 16;
    static constexpr size_t LDS_BYTES = 2 * (size_t)STAGE * 16;
    static_assert(WM % 32 == 0 && WN % 32 == 0, "wave tile must be a multiple of 32");
    static_assert(BK == 32 || BK == 64, "BK is 32 or 64");
    static_assert(PA % THREADS == 0 && PB % THREADS == 0, "panel chunks must split evenly");
};

// One BM x BN output tile (tm, tn) of batch instance bz.  The block's LDS is
// free on entry and on exit (the last k step ends with a barrier), so a
// persistent kernel may call it for tile after tile.
template <int BM, int BN, int BK, int WGM, int WGN>
__device__ __forceinline__ void x3_tile(const GemmArgs& a, uint32_t tm, uint32_t tn, uint64_t bz, u32x4* lds)
{
    using S = X3Shape<BM, BN, BK, WGM, WGN>;
    const uint64_t Kp = a.Kp;
    float* C = a.C + bz * a.c_bstride;

    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = (wave / WGN) * S::WM;
    const int wn = (wave % WGN) * S::WN;

    // this thread's load rows: row (t / CPR) + s * RA of each panel, chunk q_ld
    const int q_ld = t % S::CPR;
    const int r_ld = t / S::CPR;
    const u32x4* gAh = reinterpret_cast<const u32x4*>(a.Ah + bz * a.pa_bstride + ((uint64_t)tm * BM + r_ld) * Kp) + q_ld;
    const u32x4* gAl = reinterpret_cast<const u32x4*>(a.Al + bz * a.pa_bstride + ((uint64_t)tm * BM + r_ld) * Kp) + q_ld;
    const u32x4* gBh = reinterpret_cast<const u32x4*>(a.Bh + bz * a.pb_bstride + ((uint64_t)tn * BN + r_ld) * Kp) + q_ld;
    const u32x4* gBl = reinterpret_cast<const u32x4*>(a.Bl + bz * a.pb_bstride + ((uint64_t)tn * BN + r_ld) * Kp) + q_ld;
    const uint64_t slot_stride = (uint64_t)S::RA * Kp / 8;    // in u32x4 units
    // LDS offsets (u32x4 units, within a stage) of this thread's stores
    int lo_st[S::NA];
#pragma unroll
    for (int s = 0; s < S::NA; s++) {
        const int row = r_ld + s * S::RA;
        lo_st[s] = row * S::CPR + (q_ld ^ ((row >> S::SH) & (S::CPR - 1)));
    }
    int lo_stb[S::NB];
#pragma unroll
    for (int s = 0; s < S::NB; s++) {
        const int row = r_ld + s * S::RA;
        lo_stb[s] = row * S::CPR + (q_ld ^ ((row >> S::SH) & (S::CPR - 1)));
    }

    floatx16 acc[S::TI][S::TJ];
#pragma unroll
    for (int i = 0; i < S::TI; i++)
#pragma unroll
        for (int j = 0; j < S::TJ; j++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[i][j][r] = 0.0f;

    u32x4 sAh[S::NA], sAl[S::NA], sBh[S::NB], sBl[S::NB];

    const int nk = (int)(Kp / BK);
    // prologue: stage 0 -> LDS buffer 0
#pragma unroll
    for (int s = 0; s < S::NA; s++) { sAh[s] = gAh[s * slot_stride]; sAl[s] = gAl[s * slot_stride]; }
#pragma unroll
    for (int s = 0; s < S::NB; s++) { sBh[s] = gBh[s * slot_stride]; sBl[s] = gBl[s * slot_stride]; }
#pragma unroll
    for (int s = 0; s < S::NA; s++) { lds[lo_st[s]] = sAh[s]; lds[S::PA + lo_st[s]] = sAl[s]; }
#pragma unroll
    for (int s = 0; s < S::NB; s++) { lds[2 * S::PA + lo_stb[s]] = sBh[s]; lds[2 * S::PA + S::PB + lo_stb[s]] = sBl[s]; }
    __syncthreads();

    // fragment read offsets (u32x4 units) for each A / B tile row of this lane
    int offA[S::TI][S::KB], offB[S::TJ][S::KB];
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

    for (int kt = 0; kt < nk; kt++) {
        const int cur = (kt & 1) * S::STAGE;
        const bool more = kt + 1 < nk;
        if (more) {
            const uint64_t ko = (uint64_t)(kt + 1) * (BK / 8);
#pragma unroll
            for (int s = 0; s < S::NA; s++) { sAh[s] = gAh[s * slot_stride + ko]; sAl[s] = gAl[s * slot_stride + ko]; }
#pragma unroll
            for (int s = 0; s < S::NB; s++) { sBh[s] = gBh[s * slot_stride + ko]; sBl[s] = gBl[s * slot_stride + ko]; }
        }
#pragma unroll
        for (int kb = 0; kb < S::KB; kb++) {
            bf16x8 ah[S::TI], al[S::TI], bh[S::TJ], bl[S::TJ];
#pragma unroll
            for (int i = 0; i < S::TI; i++) {
                ah[i] = as_bf16x8(lds[cur + offA[i][kb]]);
                al[i] = as_bf16x8(lds[cur + S::PA + offA[i][kb]]);
            }
#pragma unroll
            for (int j = 0; j < S::TJ; j++) {
                bh[j] = as_bf16x8(lds[cur + offB[j][kb]]);
                bl[j] = as_bf16x8(lds[cur + S::PB + offB[j][kb]]);
            }
            // three passes over the tiles: consecutive MFMAs write different accumulators
#pragma unroll
            for (int i = 0; i < S::TI; i++)
#pragma unroll
                for (int j = 0; j < S::TJ; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
#pragma unroll
            for (int i = 0; i < S::TI; i++)
#pragma unroll
                for (int j = 0; j < S::TJ; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
            for (int i = 0; i < S::TI; i++)
#pragma unroll
                for (int j = 0; j < S::TJ; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
        }
        if (more) {
            const int nxt = ((kt + 1) & 1) * S::STAGE;
#pragma unroll
            for (int s = 0; s < S::NA; s++) { lds[nxt + lo_st[s]] = sAh[s]; lds[nxt + S::PA + lo_st[s]] = sAl[s]; }
#pragma unroll
            for (int s = 0; s < S::NB; s++) {
                lds[nxt + 2 * S::PA + lo_stb[s]] = sBh[s];
                lds[nxt + 2 * S::PA + S::PB + lo_stb[s]] = sBl[s];
            }
        }
        __syncthreads();
    }

    // 32x32x16 output layout: lane -> column lane%32; register r -> row 8*(r/4) + 4*(lane/32) + r%4
    const uint64_t row_base = (uint64_t)tm * BM + wm;
    const uint64_t col_base = (uint64_t)tn * BN + wn;
#pragma unroll
    for (int i = 0; i < S::TI; i++)
#pragma unroll
        for (int j = 0; j < S::TJ; j++) {
            const uint64_t col = col_base + j * 32 + (lane & 31);
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const uint64_t row = row_base + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                if (row < a.M && col < a.N) store_c(C, row * a.ldc + col, acc[i][j][r], a);
            }
        }
}

// The autotuner's search space for the split-BF16x3 family, in variant-id
// order: X(BM, BN, BK, waves_m, waves_n, min blocks per CU).
#define CHAI_X3_VARIANTS(X)                                                   \
    X(256, 256, 32, 2, 4, 1)   /* 0: 8 waves of 128x64               */      \
    X(256, 256, 32, 2, 2, 1)   /* 1: 4 waves of 128x128              */      \
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
