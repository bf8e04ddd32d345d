// gemm_internal.hpp -- kernel argument block and internal launchers of the
// GEMM family (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "../../include/chai_mi355x.h"

namespace chai_hip {

struct GemmArgs {
    // split-bf16 panels in k blocks of 32: element (r, k) of a panel with Rp
    // rows at (k / 32) * Rp * 32 + r * 32 + k % 32 (A: Rp = Mp, B^T: Rp = Np)
    const uint16_t *Ah, *Al, *Bh, *Bl;
    uint64_t a_kblk, b_kblk;    // elements per k block: Mp * 32, Np * 32
    // optional, per 64 x 64 split block (row block major, Kp / 64 per row
    // block): 1 if any lo value of the block is nonzero.  A tile whose whole
    // Al (Bl) panel is zero skips the products with it (null: never skip).
    const uint8_t *a_nz, *b_nz;
    uint64_t a_nz_bstride, b_nz_bstride;
    // packed operands: FMA kernel A [Mp][Kp], B [Kp][Np]; FP64 kernel A^T
    // [Kp][Mp], B [Kp][Np] (both k-major)
    const float *Af, *Bf;
    const double *Ad, *Bd;
    float* C;
    double* Cd;
    uint64_t Kp, Np, Mp;
    uint64_t lda, ldb, K;       // f32 A / B row strides and depth, for the tile fed straight from them
    uint64_t pa_bstride, pb_bstride, c_bstride;   // elements between batch instances
    uint64_t ldc, M, N;
    uint32_t tiles_m, tiles_n;
    int general;
    double alpha, beta;
    // fused elementwise epilogue on the f32 result (after alpha / beta):
    // epi == 1: c = max(c * epi_scale + epi_shift, epi_lo), each op rounded in f32
    int epi;
    float epi_scale, epi_shift, epi_lo;
};

// variant < 0 selects the current default (set by the tuner) for the mode.
int gemm_launch(const chai_gemm_desc& d, const void* A, const void* B, void* C, int mode,
                int variant, hipStream_t st, const chai_epilogue* epi = nullptr);
int gemm_num_variants(int mode);
int gemm_variant_params(int mode, int vid, int* params);   // BM, BN, BK, WGM, WGN
void gemm_set_default_variant(int mode, int vid);
// best variant recorded by the tuner for this exogenous key, or -1
int tuned_variant(const chai_gemm_desc& d, int mode);
// Padded rows of op(A) (columns of op(B)) the split-BF16x3 tile addresses
// with its 32-bit stage-copy offsets.
constexpr uint64_t kX3MaxRows = 1ull << 25;
// The family CHAI_GEMM_AUTO resolves to for a descriptor (gemm_api.hip).
int gemm_auto_mode(const chai_gemm_desc& d);
// integer x integer products with the interpreter's semantics (igemm.hip)
int igemm_launch(const chai_gemm_desc& d, const void* A, const void* B, void* C, hipStream_t st);
// the matrix-vector family (gemv.hip): variants, the untuned choice, launch
int gemv_num_variants();
int gemv_variant_params(int vid, int* p);
int gemv_default_variant(const chai_gemm_desc& d);
int gemv_launch(const chai_gemm_desc& d, const void* A, const void* x, void* y, int variant, hipStream_t st);

}  // namespace chai_hip
