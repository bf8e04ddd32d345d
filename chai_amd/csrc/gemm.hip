// gemm.hip -- the GEMM family behind chai_hip_gemm.
//
// Replaces the Evergreen "MM" OpenCL generator (autotune/EvergreenMatmulMM.cpp:
// 304-384 parameters, 1077-1630 kernel text) and its enqueue path
// (enqueue/EnqueueMatmul.cpp:24-161, enqueue/EnqueueMatmulAuto.cpp:21-402).
// Semantics kept: C = op(A) op(B) or C = alpha op(A) op(B) + beta C, row-major,
// op() folding a transpose, accumulation precision max(precA, precB), batches of
// independent instances with optional shared A or B (sameDataA/B).
//
// Kernels (all gfx950):
//   split_kernel    : f32/i32/u32 operand -> bf16 hi/lo panels in k blocks of 32,
//                     zero padded to the tile grid (an HBM-bound pre-pass).
//   gemm_bf16x3     : 256x256x32 tiles, 8 waves (2 x 4), each wave 128x64 built
//                     from 32x32x16 bf16 MFMAs; three products hi*lo + lo*hi +
//                     hi*hi accumulate in FP32; double-buffered XOR-swizzled LDS.
//   sgemm_fma       : plain FP32 VALU FMA, BM x BN x 16 tiles (BM, BN in {64, 128}),
//                     (BM/16) x (BN/16) outputs per thread.
//   dgemm_mfma      : FP64 with 16x16x4 f64 MFMAs, 64 x 64 per wave, three tile
//                     shapes (128x128, 256x128, 128x256), stages copied into LDS directly.
#include "common.hpp"
#include "gemm_device.hpp"
#include "gemm_internal.hpp"
#include "gemm_split.hpp"
#include "gemm_x3.hpp"
#include "gemm_x6.hpp"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <type_traits>

namespace chai_hip {

// =====================================================================
// split / pack pre-pass
// =====================================================================

// Block grid of the split kernels: consecutive blocks walk the source's
// contiguous index, k for a K-contiguous source (grid (Kp / 64, Rp / 64, nb)),
// r for a transposed one (grid (Rp / 64, Kp / 64, nb)); for 8192^2 f32 the
// transposed split takes 0.104 ms in r order against 0.119 ms in k order.
template <bool TRANS>
__device__ __forceinline__ void split_block(uint32_t& rb, uint32_t& kb, uint32_t& nkb)
{
    rb = TRANS ? blockIdx.x : blockIdx.y;
    kb = TRANS ? blockIdx.y : blockIdx.x;
    nkb = TRANS ? gridDim.y : gridDim.x;
}

template <typename T, bool TRANS>
// Block (rb, kb, z) splits rows [64 rb, +64) x k [64 kb, +64) of instance z
// and writes nz[(z * (Rp / 64) + rb) * (Kp / 64) + kb] = 1 if any of its lo
// values is nonzero (the GEMM tile skips products with all-zero lo panels).
__global__ __launch_bounds__(256) void split_kernel(const T* __restrict__ src, uint64_t ld,
                                                    uint64_t R, uint64_t K, uint64_t Rp, uint64_t Kp,
                                                    uint64_t src_bstride,
                                                    uint16_t* __restrict__ hi,
                                                    uint16_t* __restrict__ lo, uint8_t* __restrict__ nz)
{
    __shared__ float tile[64][65];
    uint32_t rb, kb, nkb;
    split_block<TRANS>(rb, kb, nkb);
    const uint64_t dst_bstride = Rp * Kp;
    const uint32_t lo_or =
        split_tile<T, TRANS>(src + (uint64_t)blockIdx.z * src_bstride, ld, R, K, Kp, Rp,
                             hi + (uint64_t)blockIdx.z * dst_bstride, lo + (uint64_t)blockIdx.z * dst_bstride,
                             (uint64_t)rb * 64, (uint64_t)kb * 64, tile);
    const int any = __syncthreads_or(lo_or != 0u);
    if (threadIdx.x == 0)
        nz[((uint64_t)blockIdx.z * (Rp / 64) + rb) * nkb + kb] = (uint8_t)(any != 0);
}

// The three-piece split of the BF16x6 family: p0 = hi, p1 = lo, p2 = the
// third piece, each in the k-blocked panel layout, instances dst_bstride
// apart; nz as split_kernel's (of the lo panel).
template <typename T, bool TRANS>
__global__ __launch_bounds__(256) void split3_kernel(const T* __restrict__ src, uint64_t ld, uint64_t R,
                                                     uint64_t K, uint64_t Rp, uint64_t Kp, uint64_t src_bstride,
                                                     uint64_t dst_bstride, uint16_t* __restrict__ p0,
                                                     uint16_t* __restrict__ p1, uint16_t* __restrict__ p2,
                                                     uint8_t* __restrict__ nz)
{
    __shared__ float tile[64][65];
    const uint64_t z = blockIdx.z;
    uint32_t rb, kb, nkb;
    split_block<TRANS>(rb, kb, nkb);
    const uint32_t lo_or =
        split_tile<T, TRANS, true>(src + z * src_bstride, ld, R, K, Kp, Rp, p0 + z * dst_bstride,
                                   p1 + z * dst_bstride, (uint64_t)rb * 64, (uint64_t)kb * 64,
                                   tile, p2 + z * dst_bstride);
    const int any = __syncthreads_or(lo_or != 0u);
    if (threadIdx.x == 0)
        nz[(z * (Rp / 64) + rb) * nkb + kb] = (uint8_t)(any != 0);
}

// Pack op(X) into a dense zero-padded [Rp][Cp] array of TO (element (r, c) of
// op(X) at src[r*ld + c] or, transposed, src[c*ld + r]).  A block covers 64 x 64
// elements; each thread writes 16 consecutive c of one row as 4-element
// vectors.  A row-major source is copied straight through registers (4-element
// vector loads when aligned); a transposed one goes through a padded LDS tile,
// read 4 consecutive r per lane.
template <typename T, int N>
__device__ __forceinline__ void load_vec(const T* p, T* v)
{
    typedef T TN __attribute__((ext_vector_type(N)));
    const TN x = *reinterpret_cast<const TN*>(p);
#pragma unroll
    for (int i = 0; i < N; i++) v[i] = x[i];
}

template <typename TO>
__device__ __forceinline__ void store16(TO* p, const TO* v, uint64_t n)
{
    typedef TO TO4 __attribute__((ext_vector_type(4)));
    if (n == 16 && ((uintptr_t)p & (4 * sizeof(TO) - 1)) == 0) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            TO4 x = {v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]};
            reinterpret_cast<TO4*>(p)[q] = x;
        }
    } else {
        for (uint64_t j = 0; j < n; j++) p[j] = v[j];
    }
}

template <typename T, typename TO, bool TRANS>
__global__ __launch_bounds__(256) void pack_kernel(const T* __restrict__ src, uint64_t ld, uint64_t R,
                                                   uint64_t Cn, uint64_t Rp, uint64_t Cp, uint64_t src_bstride,
                                                   uint64_t dst_bstride, TO* __restrict__ dst)
{
    __shared__ TO tile[64][65];
    const uint64_t r0 = (uint64_t)blockIdx.y * 64, c0 = (uint64_t)blockIdx.x * 64;
    src += (uint64_t)blockIdx.z * src_bstride;
    dst += (uint64_t)blockIdx.z * dst_bstride;
    const int t = threadIdx.x;
    const bool vec = ((ld & 3) == 0) && (((uintptr_t)src & (4 * sizeof(T) - 1)) == 0);
    const int rr = t >> 2, cq = (t & 3) * 16;     // this thread's output row and first column
    const uint64_t r = r0 + rr, cb = c0 + cq;
    TO v[16];
    if (!TRANS) {
        if (vec && r < R && cb + 16 <= Cn) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                T x[4];
                load_vec<T, 4>(src + r * ld + cb + 4 * q, x);
#pragma unroll
                for (int i = 0; i < 4; i++) v[4 * q + i] = (TO)x[i];
            }
        } else {
#pragma unroll
            for (int j = 0; j < 16; j++) v[j] = (r < R && cb + j < Cn) ? (TO)src[r * ld + cb + j] : (TO)0;
        }
    } else {
        // source (r, c) at src[c * ld + r]: 16 lanes read 64 consecutive r of one c
        const int r4 = (t & 15) * 4, cc0 = t >> 4;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int cc = cc0 + 16 * i;
            const uint64_t c = c0 + cc;
            if (vec && r0 + r4 + 4 <= R && c < Cn) {
                T x[4];
                load_vec<T, 4>(src + c * ld + r0 + r4, x);
#pragma unroll
                for (int j = 0; j < 4; j++) tile[cc][r4 + j] = (TO)x[j];
            } else {
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const uint64_t rj = r0 + r4 + j;
                    tile[cc][r4 + j] = (rj < R && c < Cn) ? (TO)src[c * ld + rj] : (TO)0;
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 16; j++) v[j] = tile[cq + j][rr];
    }
    if (r < Rp && cb < Cp) store16(dst + r * Cp + cb, v, cb + 16 <= Cp ? 16 : Cp - cb);
}

// =====================================================================
// split-bf16x3 MFMA GEMM: tile variants of gemm_x3_kernel (gemm_x3.hpp)
// =====================================================================

struct X3Variant {
    int BM, BN, BK, WGM, WGN;
    size_t lds;
    int threads;
    void (*kernel)(GemmArgs);
};

#define CHAI_X3V(BM, BN, BK, WGM, WGN, MINB)                                              \
    {BM, BN, BK, WGM, WGN, X3Shape<BM, BN, BK, WGM, WGN>::LDS_BYTES, WGM * WGN * 64,        \
     gemm_x3_kernel<BM, BN, BK, WGM, WGN, MINB>},

// index = variant id
static const X3Variant kX3Variants[] = {CHAI_X3_VARIANTS(CHAI_X3V)};
static constexpr int kNumX3 = sizeof(kX3Variants) / sizeof(kX3Variants[0]);

#define CHAI_X6V(BM, BN, BK, WGM, WGN, MINB)                                              \
    {BM, BN, BK, WGM, WGN, X6Shape<BM, BN, BK, WGM, WGN>::LDS_BYTES, WGM * WGN * 64,        \
     gemm_x6_kernel<BM, BN, BK, WGM, WGN, MINB>},
static const X3Variant kX6Variants[] = {CHAI_X6_VARIANTS(CHAI_X6V)};
static constexpr int kNumX6 = sizeof(kX6Variants) / sizeof(kX6Variants[0]);
static int g_x3_default = 0;

// =====================================================================
// plain FP32 VALU FMA GEMM
// =====================================================================

constexpr int F_BK = 16;

// A packed [Mp][Kp] f32, B packed [Kp][Np] f32 (both zero padded).  256
// threads as 16 x 16; each owns TM = BM / 16 rows and TN = BN / 16 columns of
// the BM x BN tile, in groups of 4 spaced 64 apart.  Every output is one fmaf
// chain over k in order, so all tile shapes give the same bits.
template <int BM, int BN, int MINB>
__global__ __launch_bounds__(256, MINB) void sgemm_fma(GemmArgs a)
{
    constexpr int TM = BM / 16, TN = BN / 16;
    constexpr int LA = BM / 64, LB = BN / 64;        // 16-byte loads per thread per stage
    constexpr int TPR = BN / 4, RPP = 256 / TPR;     // B: threads per k row, k rows per pass
    static_assert(TM % 4 == 0 && TN % 4 == 0 && LA >= 1 && LB >= 1, "tile shape");
    __shared__ float As[2][F_BK][BM + 4];
    __shared__ float Bs[2][F_BK][BN + 4];
    uint32_t tm, tn;
    tile_coords<8>(blockIdx.x, a.tiles_m, a.tiles_n, tm, tn);
    const uint64_t bz = blockIdx.z;
    const uint64_t Kp = a.Kp, Np = a.Np;
    const float* A = a.Af + bz * a.pa_bstride + (uint64_t)tm * BM * Kp;
    const float* B = a.Bf + bz * a.pb_bstride + (uint64_t)tn * BN;
    float* C = a.C + bz * a.c_bstride;
    const int t = threadIdx.x;
    const int tx = t & 15, ty = t >> 4;

    float acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++) acc[i][j] = 0.0f;

    // A tile BM x 16: thread loads rows (t >> 2) + 64 h, 4 consecutive k at (t & 3) * 4
    // B tile 16 x BN: thread loads k rows t / TPR + RPP h, 4 consecutive n at (t % TPR) * 4
    f32x4 ra[LA], rb[LB];
    auto gload = [&](int kt) {
        const uint64_t k0 = (uint64_t)kt * F_BK;
#pragma unroll
        for (int h = 0; h < LA; h++) {
            const int row = (t >> 2) + h * 64;
            ra[h] = *reinterpret_cast<const f32x4*>(A + (uint64_t)row * Kp + k0 + (t & 3) * 4);
        }
#pragma unroll
        for (int h = 0; h < LB; h++) {
            const int kr = t / TPR + h * RPP;
            rb[h] = *reinterpret_cast<const f32x4*>(B + (k0 + kr) * Np + (t % TPR) * 4);
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int h = 0; h < LA; h++) {
            const int row = (t >> 2) + h * 64, kc = (t & 3) * 4;
            As[buf][kc + 0][row] = ra[h].x;
            As[buf][kc + 1][row] = ra[h].y;
            As[buf][kc + 2][row] = ra[h].z;
            As[buf][kc + 3][row] = ra[h].w;
        }
#pragma unroll
        for (int h = 0; h < LB; h++) {
            const int kr = t / TPR + h * RPP;
            *reinterpret_cast<f32x4*>(&Bs[buf][kr][(t % TPR) * 4]) = rb[h];
        }
    };
    const int nk = (int)(Kp / F_BK);
    gload(0);
    lstore(0);
    __syncthreads();
    for (int kt = 0; kt < nk; kt++) {
        const int buf = kt & 1;
        if (kt + 1 < nk) gload(kt + 1);
#pragma unroll 2
        for (int k = 0; k < F_BK; k++) {
            float av[TM], bv[TN];
#pragma unroll
            for (int g = 0; g < TM / 4; g++) {
                const f32x4 v = *reinterpret_cast<const f32x4*>(&As[buf][k][ty * 4 + g * 64]);
                av[4 * g] = v.x; av[4 * g + 1] = v.y; av[4 * g + 2] = v.z; av[4 * g + 3] = v.w;
            }
#pragma unroll
            for (int g = 0; g < TN / 4; g++) {
                const f32x4 v = *reinterpret_cast<const f32x4*>(&Bs[buf][k][tx * 4 + g * 64]);
                bv[4 * g] = v.x; bv[4 * g + 1] = v.y; bv[4 * g + 2] = v.z; bv[4 * g + 3] = v.w;
            }
#pragma unroll
            for (int i = 0; i < TM; i++)
#pragma unroll
                for (int j = 0; j < TN; j++) acc[i][j] = __builtin_fmaf(av[i], bv[j], acc[i][j]);
        }
        if (kt + 1 < nk) lstore(buf ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < TM; i++) {
        const uint64_t row = (uint64_t)tm * BM + ty * 4 + (i & 3) + (i >> 2) * 64;
        if (row >= a.M) continue;
#pragma unroll
        for (int j = 0; j < TN; j++) {
            const uint64_t col = (uint64_t)tn * BN + tx * 4 + (j & 3) + (j >> 2) * 64;
            if (col < a.N) store_c(C, row * a.ldc + col, acc[i][j], a);
        }
    }
}

struct FVariant {
    int BM, BN;
    void (*kernel)(GemmArgs);
};
// FP32-FMA tile variants (the tuner's search space for mode FMA): BM x BN
// with the resident blocks per CU the register budget is sized for
static const FVariant kFVariants[] = {{128, 128, sgemm_fma<128, 128, 2>}, {64, 128, sgemm_fma<64, 128, 3>},
                                      {128, 64, sgemm_fma<128, 64, 3>}, {64, 64, sgemm_fma<64, 64, 4>}};
static constexpr int kNumF = sizeof(kFVariants) / sizeof(kFVariants[0]);

// =====================================================================
// FP64 MFMA GEMM
// =====================================================================

constexpr int D_BK = 8;      // k rows per LDS stage
constexpr int D_PAD = 16;    // doubles after each LDS row: rows k and k + 1 start 128 B apart mod 256 B

// A^T packed [Kp][Mp] f64 and B packed [Kp][Np] f64 (k-major, zero padded).
// Block tile BM x BN, WGM x WGN waves of 64 x 64 (4 x 4 v_mfma_f64_16x16x4f64
// accumulators).  Each stage is 8 k rows of both panels, copied straight into
// LDS one 1 KiB row piece per wave issue (global_load_lds_dwordx4); the padded
// row pitch makes the fragment reads (16 lanes per row, 4 rows) conflict free.
template <int BM, int BN, int WGM, int WGN, int MINB>
__global__ __launch_bounds__(WGM * WGN * 64, MINB) void dgemm_mfma(GemmArgs a)
{
    constexpr int WAVES = WGM * WGN;
    constexpr int PA = BM + D_PAD, PB = BN + D_PAD;          // LDS row pitch (doubles)
    constexpr int STAGE = D_BK * (PA + PB);                   // doubles per stage
    constexpr int RA = BM / 128, RB = BN / 128;               // 1 KiB pieces per k row
    constexpr int NI = D_BK * (RA + RB), IW = NI / WAVES;     // issues per stage, per wave
    static_assert(BM % 128 == 0 && BN % 128 == 0 && NI % WAVES == 0, "whole issues per wave");
    static_assert(BM / WGM == 64 && BN / WGN == 64, "64 x 64 per wave");
    extern __shared__ double dlds[];
    uint32_t tm, tn;
    tile_coords<8>(blockIdx.x, a.tiles_m, a.tiles_n, tm, tn);
    const uint64_t bz = blockIdx.z;
    const uint64_t Kp = a.Kp, Np = a.Np, Mp = a.Mp;
    const double* At = a.Ad + bz * a.pa_bstride + (uint64_t)tm * BM;
    const double* Bp = a.Bd + bz * a.pb_bstride + (uint64_t)tn * BN;
    double* C = a.Cd + bz * a.c_bstride;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = (wave / WGN) * 64, wn = (wave % WGN) * 64;

    doublex4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = (doublex4){0.0, 0.0, 0.0, 0.0};

    const int nk = (int)(Kp / D_BK);
#define DGEMM_ISSUE(KT, ST)                                                                          \
    {                                                                                                \
        _Pragma("unroll") for (int s = 0; s < IW; s++) {                                             \
            const int j = wave * IW + s;   /* wave-uniform */                                        \
            if (j < D_BK * RA) {                                                                     \
                const int k = j / RA, h = j % RA;                                                    \
                lds_dma16(At + ((uint64_t)(KT) * D_BK + k) * Mp + h * 128 + lane * 2,                \
                          (ST) + k * PA + h * 128);                                                  \
            } else {                                                                                 \
                const int k = (j - D_BK * RA) / RB, h = (j - D_BK * RA) % RB;                        \
                lds_dma16(Bp + ((uint64_t)(KT) * D_BK + k) * Np + h * 128 + lane * 2,                \
                          (ST) + D_BK * PA + k * PB + h * 128);                                      \
            }                                                                                        \
        }                                                                                            \
    }
    DGEMM_ISSUE(0, dlds)
    wait_lds_dma();
    __syncthreads();
    for (int kt = 0; kt < nk; kt++) {
        const double* cur = dlds + (kt & 1) * STAGE;
        if (kt + 1 < nk) DGEMM_ISSUE(kt + 1, dlds + ((kt + 1) & 1) * STAGE)
#pragma unroll
        for (int ks = 0; ks < D_BK / 4; ks++) {
            const int k = ks * 4 + (lane >> 4);
            double av[4], bv[4];
#pragma unroll
            for (int i = 0; i < 4; i++) av[i] = cur[k * PA + wm + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < 4; j++) bv[j] = cur[D_BK * PA + k * PB + wn + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bv[j], acc[i][j], 0, 0, 0);
        }
        wait_lds_dma();
        __syncthreads();
    }
#undef DGEMM_ISSUE
    // f64 16x16x4 output layout (measured on gfx950): lane -> column lane%16;
    // register r -> row 4*r + lane/16 (A row i supplied by lane i + 16k).
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint64_t col = (uint64_t)tn * BN + wn + j * 16 + (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const uint64_t row = (uint64_t)tm * BM + wm + i * 16 + 4 * r + (lane >> 4);
                if (row < a.M && col < a.N) {
                    const uint64_t idx = row * a.ldc + col;
                    const double v = acc[i][j][r];
                    C[idx] = a.general ? a.alpha * v + a.beta * C[idx] : v;
                }
            }
        }
}

struct DVariant {
    int BM, BN, WGM, WGN;
    size_t lds;
    void (*kernel)(GemmArgs);
};
#define CHAI_DV(BM, BN, WGM, WGN, MINB) \
    {BM, BN, WGM, WGN, 2 * (size_t)D_BK * (BM + D_PAD + BN + D_PAD) * 8, dgemm_mfma<BM, BN, WGM, WGN, MINB>}
// FP64 tile variants (the tuner's search space for mode F64), with the
// resident blocks per CU the register budget is sized for
static const DVariant kDVariants[] = {CHAI_DV(128, 128, 2, 2, 2), CHAI_DV(256, 128, 4, 2, 1),
                                      CHAI_DV(128, 256, 2, 4, 1)};
static constexpr int kNumD = sizeof(kDVariants) / sizeof(kDVariants[0]);

// =====================================================================
// host side
// =====================================================================

static uint64_t round_up(uint64_t x, uint64_t m) { return (x + m - 1) / m * m; }

// ---- optional per-kernel timing of the calling thread's GEMMs: a ring of
// event triples (before the pre-pass, before the core kernel, after it), one
// slot per call, so back-to-back calls are timed without a host wait between
static volatile int g_profile = 0;
constexpr int kProfSlots = 64;
struct ProfRing {
    hipEvent_t e[kProfSlots][3] = {};
    int cur = -1;              // slot of the call in progress / last call
    uint64_t done = 0;         // calls whose three events were recorded
    ~ProfRing()
    {
        for (auto& s : e)
            for (auto& x : s)
                if (x) (void)hipEventDestroy(x);
    }
};
static thread_local ProfRing t_prof;

static void prof_mark(int i, hipStream_t st)
{
    if (!g_profile) return;
    if (i == 0) t_prof.cur = (t_prof.cur + 1) % kProfSlots;
    if (t_prof.cur < 0) return;
    hipEvent_t& ev = t_prof.e[t_prof.cur][i];
    if (!ev && hipEventCreate(&ev) != hipSuccess) { ev = nullptr; return; }
    (void)hipEventRecord(ev, st);
    if (i == 2) t_prof.done++;
}

// Kp and Rp are multiples of 64: the grid is exactly (Kp / 64, Rp / 64, nb),
// or (Rp / 64, Kp / 64, nb) for a transposed source (split_block), and nz
// receives (Rp / 64) * (Kp / 64) flags per instance.
template <typename T>
static void launch_split(const void* src, uint64_t ld, bool trans, uint64_t R, uint64_t K,
                         uint64_t Rp, uint64_t Kp, uint64_t src_bstride, uint64_t nb,
                         uint16_t* hi, uint16_t* lo, uint8_t* nz, hipStream_t st)
{
    if (trans)
        split_kernel<T, true><<<dim3((unsigned)(Rp / 64), (unsigned)(Kp / 64), (unsigned)nb), 256, 0, st>>>(
            (const T*)src, ld, R, K, Rp, Kp, src_bstride, hi, lo, nz);
    else
        split_kernel<T, false><<<dim3((unsigned)(Kp / 64), (unsigned)(Rp / 64), (unsigned)nb), 256, 0, st>>>(
            (const T*)src, ld, R, K, Rp, Kp, src_bstride, hi, lo, nz);
}

static void split_any(int prec, const void* src, uint64_t ld, bool trans, uint64_t R, uint64_t K,
                      uint64_t Rp, uint64_t Kp, uint64_t src_bstride, uint64_t nb, uint16_t* hi,
                      uint16_t* lo, uint8_t* nz, hipStream_t st)
{
    switch (prec) {
        case CHAI_F32: launch_split<float>(src, ld, trans, R, K, Rp, Kp, src_bstride, nb, hi, lo, nz, st); break;
        case CHAI_I32: launch_split<int32_t>(src, ld, trans, R, K, Rp, Kp, src_bstride, nb, hi, lo, nz, st); break;
        default: launch_split<uint32_t>(src, ld, trans, R, K, Rp, Kp, src_bstride, nb, hi, lo, nz, st); break;
    }
}

template <typename T>
static void launch_split3(const void* src, uint64_t ld, bool trans, uint64_t R, uint64_t K, uint64_t Rp,
                          uint64_t Kp, uint64_t src_bstride, uint64_t dst_bstride, uint64_t nb, uint16_t* p0,
                          uint16_t* p1, uint16_t* p2, uint8_t* nz, hipStream_t st)
{
    if (trans)
        split3_kernel<T, true><<<dim3((unsigned)(Rp / 64), (unsigned)(Kp / 64), (unsigned)nb), 256, 0, st>>>(
            (const T*)src, ld, R, K, Rp, Kp, src_bstride, dst_bstride, p0, p1, p2, nz);
    else
        split3_kernel<T, false><<<dim3((unsigned)(Kp / 64), (unsigned)(Rp / 64), (unsigned)nb), 256, 0, st>>>(
            (const T*)src, ld, R, K, Rp, Kp, src_bstride, dst_bstride, p0, p1, p2, nz);
}

static void split3_any(int prec, const void* src, uint64_t ld, bool trans, uint64_t R, uint64_t K, uint64_t Rp,
                       uint64_t Kp, uint64_t src_bstride, uint64_t dst_bstride, uint64_t nb, uint16_t* p0,
                       uint16_t* p1, uint16_t* p2, uint8_t* nz, hipStream_t st)
{
    switch (prec) {
        case CHAI_F32: launch_split3<float>(src, ld, trans, R, K, Rp, Kp, src_bstride, dst_bstride, nb, p0, p1, p2, nz, st); break;
        case CHAI_I32: launch_split3<int32_t>(src, ld, trans, R, K, Rp, Kp, src_bstride, dst_bstride, nb, p0, p1, p2, nz, st); break;
        default: launch_split3<uint32_t>(src, ld, trans, R, K, Rp, Kp, src_bstride, dst_bstride, nb, p0, p1, p2, nz, st); break;
    }
}

template <typename T, typename TO>
static void launch_pack(const void* src, uint64_t ld, bool trans, uint64_t R, uint64_t Cn,
                        uint64_t Rp, uint64_t Cp, uint64_t src_bstride, uint64_t nb, TO* dst,
                        hipStream_t st)
{
    dim3 grid((unsigned)((Cp + 63) / 64), (unsigned)((Rp + 63) / 64), (unsigned)nb);
    if (trans)
        pack_kernel<T, TO, true><<<grid, 256, 0, st>>>((const T*)src, ld, R, Cn, Rp, Cp, src_bstride, Rp * Cp, dst);
    else
        pack_kernel<T, TO, false><<<grid, 256, 0, st>>>((const T*)src, ld, R, Cn, Rp, Cp, src_bstride, Rp * Cp, dst);
}

template <typename TO>
static void pack_any(int prec, const void* src, uint64_t ld, bool trans, uint64_t R, uint64_t Cn,
                     uint64_t Rp, uint64_t Cp, uint64_t src_bstride, uint64_t nb, TO* dst, hipStream_t st)
{
    switch (prec) {
        case CHAI_F32: launch_pack<float, TO>(src, ld, trans, R, Cn, Rp, Cp, src_bstride, nb, dst, st); break;
        case CHAI_F64: launch_pack<double, TO>(src, ld, trans, R, Cn, Rp, Cp, src_bstride, nb, dst, st); break;
        case CHAI_I32: launch_pack<int32_t, TO>(src, ld, trans, R, Cn, Rp, Cp, src_bstride, nb, dst, st); break;
        default: launch_pack<uint32_t, TO>(src, ld, trans, R, Cn, Rp, Cp, src_bstride, nb, dst, st); break;
    }
}

// FP64 tile without tuning: the fewest tile-rounds x tile area, i.e. the least
// MFMA time once the last partial round of resident blocks is counted.
static int f64_default_variant(uint64_t M, uint64_t N, uint64_t nb)
{
    static int slots[kNumD] = {};
    static std::mutex mu;
    std::lock_guard<std::mutex> lk(mu);
    int best = 0;
    double best_cost = 0;
    for (int v = 0; v < kNumD; v++) {
        const DVariant& d = kDVariants[v];
        if (!slots[v]) {
            int dev = 0, cus = 0, per_cu = 0;
            if (hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
                hipFuncSetAttribute((const void*)d.kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)d.lds) !=
                    hipSuccess ||
                hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, d.kernel, d.WGM * d.WGN * 64, d.lds) !=
                    hipSuccess) {
                set_error("gemm: f64 occupancy query failed");
                return -1;
            }
            slots[v] = std::max(1, cus * std::max(per_cu, 1));
        }
        const double tiles = (double)((M + d.BM - 1) / d.BM) * (double)((N + d.BN - 1) / d.BN) * (double)nb;
        const double cost = std::ceil(tiles / slots[v]) * d.BM * d.BN;
        if (v == 0 || cost < best_cost) { best = v; best_cost = cost; }
    }
    return best;
}

static void set_epilogue(GemmArgs& g, const chai_epilogue* epi)
{
    if (epi && epi->kind == CHAI_EPI_AFFINE_MAX) {
        g.epi = 1;
        g.epi_scale = epi->scale; g.epi_shift = epi->shift; g.epi_lo = epi->lo;
    }
}

int gemm_launch(const chai_gemm_desc& d, const void* A, const void* B, void* C, int mode,
                int variant, hipStream_t st, const chai_epilogue* epi)
{
    if (epi && epi->kind != CHAI_EPI_NONE && epi->kind != CHAI_EPI_AFFINE_MAX) {
        set_error("gemm: unknown epilogue kind %d", epi->kind);
        return CHAI_ERR_ARG;
    }
    if (epi && epi->kind != CHAI_EPI_NONE && mode == CHAI_GEMM_F64) {
        set_error("gemm: epilogues apply to f32 results only");
        return CHAI_ERR_ARG;
    }
    const uint64_t M = d.M, N = d.N, K = d.K, nb = d.batch ? d.batch : 1;
    const uint64_t a_rows = d.trans_a ? K : M;   // stored layout of A
    (void)a_rows;
    if (mode == CHAI_GEMM_BF16X3) {
        if (d.prec_c != CHAI_F32) { set_error("gemm: bf16x3 writes f32 C"); return CHAI_ERR_ARG; }
        const int vid = variant < 0 ? g_x3_default : variant;
        if (vid >= kNumX3) { set_error("gemm: bf16x3 variant %d out of range", vid); return CHAI_ERR_ARG; }
        const X3Variant& v = kX3Variants[vid];
        const uint64_t Mp = round_up(M, v.BM), Np = round_up(N, v.BN), Kp = round_up(K, 64);
        // the tile's stage copies address a panel's k block pair with 32-bit
        // byte offsets (2 x 32 x Mp bytes per k block)
        if (Mp > kX3MaxRows || Np > kX3MaxRows) {
            set_error("gemm: bf16x3 takes at most %llu rows of op(A) and columns of op(B) per call",
                      (unsigned long long)kX3MaxRows);
            return CHAI_ERR_ARG;
        }
        const uint64_t na = d.same_a ? 1 : nb, nbb = d.same_b ? 1 : nb;
        const size_t panel_bytes = 2 * sizeof(uint16_t) * (Mp * Kp * na + Np * Kp * nbb);
        const uint64_t fa = (Mp / 64) * (Kp / 64), fb = (Np / 64) * (Kp / 64);   // split flags per instance
        const size_t bytes = panel_bytes + fa * na + fb * nbb;
        uint16_t* ws = nullptr;
        CHAI_HIP_TRY(hipMallocAsync((void**)&ws, bytes, st));
        uint16_t* Ah = ws;
        uint16_t* Al = Ah + Mp * Kp * na;
        uint16_t* Bh = Al + Mp * Kp * na;
        uint16_t* Bl = Bh + Np * Kp * nbb;
        uint8_t* nz_a = (uint8_t*)ws + panel_bytes;
        uint8_t* nz_b = nz_a + fa * na;
        prof_mark(0, st);
        // op(A): element (m,k) at A[m*lda+k] or, transposed, A[k*lda+m]
        split_any(d.prec_a, A, d.lda, d.trans_a != 0, M, K, Mp, Kp, d.stride_a, na, Ah, Al, nz_a, st);
        // B^T panels: element (n,k) = op(B)[k][n] at B[k*ldb+n] (transposed read) or B[n*ldb+k]
        split_any(d.prec_b, B, d.ldb, d.trans_b == 0, N, K, Np, Kp, d.stride_b, nbb, Bh, Bl, nz_b, st);
        prof_mark(1, st);
        GemmArgs g{};
        g.Ah = Ah; g.Al = Al; g.Bh = Bh; g.Bl = Bl;
        g.a_kblk = Mp * 32; g.b_kblk = Np * 32;
        g.a_nz = nz_a; g.b_nz = nz_b;
        g.a_nz_bstride = d.same_a ? 0 : fa;
        g.b_nz_bstride = d.same_b ? 0 : fb;
        g.Kp = Kp; g.Np = Np;
        g.pa_bstride = d.same_a ? 0 : Mp * Kp;
        g.pb_bstride = d.same_b ? 0 : Np * Kp;
        g.C = (float*)C; g.c_bstride = d.stride_c; g.ldc = d.ldc;
        g.M = M; g.N = N;
        g.tiles_m = (uint32_t)(Mp / v.BM); g.tiles_n = (uint32_t)(Np / v.BN);
        g.general = d.general; g.alpha = d.alpha; g.beta = d.beta;
        set_epilogue(g, epi);
        static bool attr_set[kNumX3] = {};
        if (!attr_set[vid]) {
            CHAI_HIP_TRY(hipFuncSetAttribute((const void*)v.kernel,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)v.lds));
            attr_set[vid] = true;
        }
        dim3 grid(g.tiles_m * g.tiles_n, 1, (unsigned)nb);
        hipLaunchKernelGGL(v.kernel, grid, dim3(v.threads), v.lds, st, g);
        hipError_t e = hipGetLastError();
        prof_mark(2, st);
        CHAI_HIP_TRY(hipFreeAsync(ws, st));
        if (e != hipSuccess) return hip_fail(e, "gemm_x3 launch");
        return CHAI_OK;
    }
    if (mode == CHAI_GEMM_BF16X6) {
        // Split-BF16x6: A = a0 + a1 + a2, B = b0 + b1 + b2 (three bf16 pieces,
        // all 24 bits of an f32); one pass of the six-panel tile (gemm_x6.hpp)
        // over the split3 pre-pass's panels.  The terms dropped (order >= 3)
        // are below 2^-24 |a b|.
        if (d.prec_c != CHAI_F32) { set_error("gemm: bf16x6 writes f32 C"); return CHAI_ERR_ARG; }
        const int vid = variant < 0 ? 0 : variant;
        if (vid >= kNumX6) { set_error("gemm: bf16x6 variant %d out of range", vid); return CHAI_ERR_ARG; }
        const X3Variant& v = kX6Variants[vid];
        const uint64_t Mp = round_up(M, v.BM), Np = round_up(N, v.BN), Kp = round_up(K, 64);
        const uint64_t na = d.same_a ? 1 : nb, nbb = d.same_b ? 1 : nb;
        const uint64_t sa3 = 3 * Mp * Kp, sb3 = 3 * Np * Kp;        // elements per instance
        const size_t panel_bytes = sizeof(uint16_t) * (sa3 * na + sb3 * nbb);
        const uint64_t fa = (Mp / 64) * (Kp / 64), fb = (Np / 64) * (Kp / 64);
        const size_t bytes = panel_bytes + fa * na + fb * nbb;
        uint16_t* ws = nullptr;
        CHAI_HIP_TRY(hipMallocAsync((void**)&ws, bytes, st));
        uint16_t* PA = ws;                       // per instance [a0 | a1 | a2]
        uint16_t* PB = PA + sa3 * na;            // per instance [b0 | b1 | b2]
        uint8_t* nz_a = (uint8_t*)ws + panel_bytes;
        uint8_t* nz_b = nz_a + fa * na;
        prof_mark(0, st);
        split3_any(d.prec_a, A, d.lda, d.trans_a != 0, M, K, Mp, Kp, d.stride_a, sa3, na,
                   PA, PA + Mp * Kp, PA + 2 * Mp * Kp, nz_a, st);
        split3_any(d.prec_b, B, d.ldb, d.trans_b == 0, N, K, Np, Kp, d.stride_b, sb3, nbb,
                   PB, PB + Np * Kp, PB + 2 * Np * Kp, nz_b, st);
        prof_mark(1, st);
        static bool attr6[kNumX6] = {};
        if (!attr6[vid]) {
            CHAI_HIP_TRY(hipFuncSetAttribute((const void*)v.kernel,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)v.lds));
            attr6[vid] = true;
        }
        GemmArgs g{};
        g.Ah = PA; g.Bh = PB;
        g.a_kblk = Mp * 32; g.b_kblk = Np * 32;
        g.a_nz = nz_a; g.b_nz = nz_b;
        g.a_nz_bstride = d.same_a ? 0 : fa;
        g.b_nz_bstride = d.same_b ? 0 : fb;
        g.Kp = Kp; g.Np = Np; g.Mp = Mp;
        g.pa_bstride = d.same_a ? 0 : sa3;
        g.pb_bstride = d.same_b ? 0 : sb3;
        g.C = (float*)C; g.c_bstride = d.stride_c; g.ldc = d.ldc;
        g.M = M; g.N = N;
        g.tiles_m = (uint32_t)(Mp / v.BM); g.tiles_n = (uint32_t)(Np / v.BN);
        g.general = d.general; g.alpha = d.alpha; g.beta = d.beta;
        set_epilogue(g, epi);
        dim3 grid(g.tiles_m * g.tiles_n, 1, (unsigned)nb);
        hipLaunchKernelGGL(v.kernel, grid, dim3(v.threads), v.lds, st, g);
        hipError_t e = hipGetLastError();
        prof_mark(2, st);
        CHAI_HIP_TRY(hipFreeAsync(ws, st));
        if (e != hipSuccess) return hip_fail(e, "gemm_x6 launch");
        return CHAI_OK;
    }
    if (mode == CHAI_GEMM_GEMV) {
        prof_mark(0, st);
        prof_mark(1, st);
        const int rc = gemv_launch(d, A, B, C, variant, st);
        prof_mark(2, st);
        return rc;
    }
    if (mode == CHAI_GEMM_FMA || mode == CHAI_GEMM_F64) {
        const bool f64 = mode == CHAI_GEMM_F64;
        int dv = 0;
        if (f64) {
            if (variant >= kNumD) { set_error("gemm: f64 variant %d out of range", variant); return CHAI_ERR_ARG; }
            dv = variant >= 0 ? variant : f64_default_variant(M, N, nb);
            if (dv < 0) return CHAI_ERR_HIP;
        }
        int fv = 0;
        if (!f64) {
            if (variant >= kNumF) { set_error("gemm: fma variant %d out of range", variant); return CHAI_ERR_ARG; }
            fv = variant >= 0 ? variant : 0;
        }
        const uint64_t BMx = f64 ? kDVariants[dv].BM : kFVariants[fv].BM;
        const uint64_t BNx = f64 ? kDVariants[dv].BN : kFVariants[fv].BN;
        const uint64_t BKx = f64 ? D_BK : F_BK;
        const uint64_t Mp = round_up(M, BMx), Np = round_up(N, BNx), Kp = round_up(K, 32);
        const uint64_t na = d.same_a ? 1 : nb, nbb = d.same_b ? 1 : nb;
        const size_t es = f64 ? 8 : 4;
        const size_t bytes = es * (Mp * Kp * na + Kp * Np * nbb);
        void* ws = nullptr;
        CHAI_HIP_TRY(hipMallocAsync(&ws, bytes, st));
        char* pa = (char*)ws;
        char* pb = pa + es * Mp * Kp * na;
        // FMA: A packed [Mp][Kp] (op(A) row-major); FP64: A^T packed [Kp][Mp]
        // (element (k, m) = op(A)[m][k]); B packed [Kp][Np] (op(B) row-major)
        prof_mark(0, st);
        if (f64) {
            pack_any<double>(d.prec_a, A, d.lda, d.trans_a == 0, K, M, Kp, Mp, d.stride_a, na, (double*)pa, st);
            pack_any<double>(d.prec_b, B, d.ldb, d.trans_b != 0, K, N, Kp, Np, d.stride_b, nbb, (double*)pb, st);
        } else {
            pack_any<float>(d.prec_a, A, d.lda, d.trans_a != 0, M, K, Mp, Kp, d.stride_a, na, (float*)pa, st);
            pack_any<float>(d.prec_b, B, d.ldb, d.trans_b != 0, K, N, Kp, Np, d.stride_b, nbb, (float*)pb, st);
        }
        GemmArgs g{};
        g.Kp = Kp; g.Np = Np; g.Mp = Mp;
        g.pa_bstride = d.same_a ? 0 : Mp * Kp;
        g.pb_bstride = d.same_b ? 0 : Kp * Np;
        g.c_bstride = d.stride_c; g.ldc = d.ldc;
        g.M = M; g.N = N;
        g.tiles_m = (uint32_t)(Mp / BMx); g.tiles_n = (uint32_t)(Np / BNx);
        g.general = d.general; g.alpha = d.alpha; g.beta = d.beta;
        set_epilogue(g, epi);
        dim3 grid(g.tiles_m * g.tiles_n, 1, (unsigned)nb);
        (void)BKx;
        prof_mark(1, st);
        if (f64) {
            g.Ad = (const double*)pa; g.Bd = (const double*)pb; g.Cd = (double*)C;
            const DVariant& v = kDVariants[dv];
            static bool d_attr[kNumD] = {};
            if (!d_attr[dv]) {
                CHAI_HIP_TRY(hipFuncSetAttribute((const void*)v.kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 (int)v.lds));
                d_attr[dv] = true;
            }
            hipLaunchKernelGGL(v.kernel, grid, dim3(v.WGM * v.WGN * 64), v.lds, st, g);
        } else {
            g.Af = (const float*)pa; g.Bf = (const float*)pb; g.C = (float*)C;
            hipLaunchKernelGGL(kFVariants[fv].kernel, grid, dim3(256), 0, st, g);
        }
        hipError_t e = hipGetLastError();
        prof_mark(2, st);
        CHAI_HIP_TRY(hipFreeAsync(ws, st));
        if (e != hipSuccess) return hip_fail(e, f64 ? "dgemm_mfma launch" : "sgemm_fma launch");
        return CHAI_OK;
    }
    set_error("gemm: unknown mode %d", mode);
    return CHAI_ERR_ARG;
}

}  // namespace chai_hip

using namespace chai_hip;

namespace chai_hip {

int gemm_num_variants(int mode)
{
    return mode == CHAI_GEMM_BF16X3                                 ? kNumX3
           : mode == CHAI_GEMM_BF16X6                                ? kNumX6
           : mode == CHAI_GEMM_F64                                   ? kNumD
           : mode == CHAI_GEMM_FMA                                   ? kNumF
           : mode == CHAI_GEMM_GEMV                                  ? gemv_num_variants()
                                                                     : 1;
}

int gemm_variant_params(int mode, int vid, int* p)
{
    if (mode == CHAI_GEMM_F64 && vid >= 0 && vid < kNumD) {
        const DVariant& v = kDVariants[vid];
        p[0] = v.BM; p[1] = v.BN; p[2] = D_BK; p[3] = v.WGM; p[4] = v.WGN;
        return CHAI_OK;
    }
    if (mode == CHAI_GEMM_GEMV) return gemv_variant_params(vid, p);
    if (mode == CHAI_GEMM_FMA && vid >= 0 && vid < kNumF) {
        const FVariant& v = kFVariants[vid];
        p[0] = v.BM; p[1] = v.BN; p[2] = F_BK; p[3] = 4; p[4] = 1;   // 4 waves, each 16 / 4 thread rows
        return CHAI_OK;
    }
    if (mode == CHAI_GEMM_BF16X6 && vid >= 0 && vid < kNumX6) {
        const X3Variant& v = kX6Variants[vid];
        p[0] = v.BM; p[1] = v.BN; p[2] = v.BK; p[3] = v.WGM; p[4] = v.WGN;
        return CHAI_OK;
    }
    if (mode != CHAI_GEMM_BF16X3 || vid < 0 || vid >= kNumX3) return CHAI_ERR_ARG;
    const X3Variant& v = kX3Variants[vid];
    p[0] = v.BM; p[1] = v.BN; p[2] = v.BK; p[3] = v.WGM; p[4] = v.WGN;
    return CHAI_OK;
}

void gemm_set_default_variant(int mode, int vid)
{
    if (mode == CHAI_GEMM_BF16X3 && vid >= 0 && vid < kNumX3) g_x3_default = vid;
}

}  // namespace chai_hip

extern "C" int chai_hip_profile_enable(int on)
{
    g_profile = on ? 1 : 0;
    if (on) t_prof.done = 0;   // the calling thread's ring starts over
    return CHAI_OK;
}

// durations of ring slot k (waits for its last event)
static int prof_slot(int k, float* prep_ms, float* core_ms)
{
    hipEvent_t* e = t_prof.e[k];
    if (!e[0] || !e[1] || !e[2]) { set_error("gemm timing: incomplete events"); return CHAI_ERR_STATE; }
    CHAI_HIP_TRY(hipEventSynchronize(e[2]));
    CHAI_HIP_TRY(hipEventElapsedTime(prep_ms, e[0], e[1]));
    CHAI_HIP_TRY(hipEventElapsedTime(core_ms, e[1], e[2]));
    return CHAI_OK;
}

extern "C" int chai_hip_gemm_last_timing(float* prep_ms, float* core_ms)
{
    if (t_prof.done == 0 || t_prof.cur < 0) {
        set_error("gemm_last_timing: no profiled GEMM on this thread (enable profiling first)");
        return CHAI_ERR_STATE;
    }
    float a = 0.f, b = 0.f;
    const int rc = prof_slot(t_prof.cur, &a, &b);
    if (rc != CHAI_OK) return rc;
    if (prep_ms) *prep_ms = a;
    if (core_ms) *core_ms = b;
    return CHAI_OK;
}

extern "C" int chai_hip_gemm_profile_mean(int count, float* prep_ms, float* core_ms)
{
    if (count < 1 || count > kProfSlots || (uint64_t)count > t_prof.done || t_prof.cur < 0) {
        set_error("gemm_profile_mean: %d calls asked, %llu profiled on this thread (at most %d kept)", count,
                  (unsigned long long)t_prof.done, kProfSlots);
        return CHAI_ERR_ARG;
    }
    double sa = 0, sb = 0;
    for (int i = 0; i < count; i++) {
        float a = 0.f, b = 0.f;
        const int rc = prof_slot((t_prof.cur - i + kProfSlots) % kProfSlots, &a, &b);
        if (rc != CHAI_OK) return rc;
        sa += a; sb += b;
    }
    if (prep_ms) *prep_ms = (float)(sa / count);
    if (core_ms) *core_ms = (float)(sb / count);
    return CHAI_OK;
}
