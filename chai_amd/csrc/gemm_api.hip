// gemm_api.hip -- chai_hip_gemm: argument validation and kernel-family choice.
#include "common.hpp"
#include "gemm_internal.hpp"

using namespace chai_hip;

extern "C" int chai_hip_gemm_epilogue(const chai_gemm_desc* desc, const void* A, const void* B,
                                      void* C, const chai_epilogue* epi, void* stream)
{
    if (!desc) { set_error("gemm: null descriptor"); return CHAI_ERR_ARG; }
    const chai_gemm_desc& d = *desc;
    if (!valid_prec(d.prec_a) || !valid_prec(d.prec_b) || !valid_prec(d.prec_c)) {
        set_error("gemm: bad precision (%d, %d, %d)", d.prec_a, d.prec_b, d.prec_c);
        return CHAI_ERR_ARG;
    }
    if (d.M == 0 || d.N == 0) return CHAI_OK;
    if (!A || !B || !C) { set_error("gemm: null operand"); return CHAI_ERR_ARG; }
    const size_t a_cols = d.trans_a ? d.M : d.K, b_cols = d.trans_b ? d.K : d.N;
    if (d.lda < a_cols || d.ldb < b_cols || d.ldc < d.N) {
        set_error("gemm: leading dimension smaller than the row length");
        return CHAI_ERR_ARG;
    }
    if (d.batch > 65535) { set_error("gemm: at most 65535 batch instances per call"); return CHAI_ERR_ARG; }
    const bool any_f64 = d.prec_a == CHAI_F64 || d.prec_b == CHAI_F64;
    const bool ints_only = d.prec_a <= CHAI_I32 && d.prec_b <= CHAI_I32;
    if (ints_only || d.mode == CHAI_GEMM_INT) {
        if (!ints_only || (d.mode != CHAI_GEMM_AUTO && d.mode != CHAI_GEMM_INT)) {
            set_error("gemm: integer x integer operands take CHAI_GEMM_INT (or AUTO), and CHAI_GEMM_INT takes no others");
            return CHAI_ERR_ARG;
        }
        if (epi && epi->kind != CHAI_EPI_NONE) { set_error("gemm: epilogues apply to f32 results only"); return CHAI_ERR_ARG; }
        if (d.K == 0) { set_error("gemm: K == 0"); return CHAI_ERR_ARG; }
        return igemm_launch(d, A, B, C, as_stream(stream));
    }
    if (d.mode == CHAI_GEMM_GEMV) {
        const int want_c = any_f64 ? CHAI_F64 : CHAI_F32;
        if (d.N != 1 || (!d.trans_b && d.ldb != 1 && d.K > 1) || (d.ldc != 1 && d.M > 1)) {
            set_error("gemm: CHAI_GEMM_GEMV takes one contiguous column op(B) and writes one contiguous column C");
            return CHAI_ERR_ARG;
        }
        if (d.prec_c != want_c) { set_error("gemm: CHAI_GEMM_GEMV writes the promoted precision (%d)", want_c); return CHAI_ERR_ARG; }
        if (epi && epi->kind != CHAI_EPI_NONE) { set_error("gemm: CHAI_GEMM_GEMV has no epilogue"); return CHAI_ERR_ARG; }
        if (d.K == 0) { set_error("gemm: K == 0"); return CHAI_ERR_ARG; }
        if (d.variant >= gemv_num_variants()) { set_error("gemm: variant %d out of range for GEMV", d.variant); return CHAI_ERR_ARG; }
        const int tv = d.variant >= 0 ? d.variant : tuned_variant(d, CHAI_GEMM_GEMV);
        return gemm_launch(d, A, B, C, CHAI_GEMM_GEMV, tv >= 0 ? tv : gemv_default_variant(d), as_stream(stream),
                           nullptr);
    }
    int mode = d.mode;
    if (mode == CHAI_GEMM_AUTO) mode = gemm_auto_mode(d);
    if (any_f64 && mode != CHAI_GEMM_F64) {
        set_error("gemm: f64 operands need CHAI_GEMM_F64");
        return CHAI_ERR_ARG;
    }
    if (mode == CHAI_GEMM_F64 && d.prec_c != CHAI_F64) {
        set_error("gemm: CHAI_GEMM_F64 writes an f64 C");
        return CHAI_ERR_ARG;
    }
    if ((mode == CHAI_GEMM_FMA || mode == CHAI_GEMM_BF16X3 || mode == CHAI_GEMM_BF16X6) && d.prec_c != CHAI_F32) {
        set_error("gemm: FMA / BF16X3 / BF16X6 write an f32 C");
        return CHAI_ERR_ARG;
    }
    if (d.K == 0) {
        set_error("gemm: K == 0");
        return CHAI_ERR_ARG;
    }
    if (d.variant >= gemm_num_variants(mode)) {
        set_error("gemm: variant %d out of range for mode %d", d.variant, mode);
        return CHAI_ERR_ARG;
    }
    const int variant = d.variant >= 0 ? d.variant : tuned_variant(d, mode);
    return gemm_launch(d, A, B, C, mode, variant, as_stream(stream), epi);
}

namespace chai_hip {

// The family AUTO picks: FP64 MFMA for f64 operands, the exact integer path
// for two integer operands; for f32 (and mixed int / f32) products the split-
// BF16 form whose error against FP64 stays within twice the FP32-FMA kernel's
// in max and rms (measured per K on cancelling uniform[-1, 1) data): three
// products from K = 16384 on, six below.
int gemm_auto_mode(const chai_gemm_desc& d)
{
    if (d.prec_a == CHAI_F64 || d.prec_b == CHAI_F64) return CHAI_GEMM_F64;
    if (d.prec_a <= CHAI_I32 && d.prec_b <= CHAI_I32) return CHAI_GEMM_INT;
    if (d.K < 16384) return CHAI_GEMM_BF16X6;
    // bf16x3's panels hold at most kX3MaxRows padded rows (256-row tiles)
    return d.M + 256 <= kX3MaxRows && d.N + 256 <= kX3MaxRows ? CHAI_GEMM_BF16X3 : CHAI_GEMM_BF16X6;
}

}  // namespace chai_hip

extern "C" int chai_hip_gemm(const chai_gemm_desc* desc, const void* A, const void* B, void* C,
                             void* stream)
{
    return chai_hip_gemm_epilogue(desc, A, B, C, nullptr, stream);
}

extern "C" int chai_hip_gemm_num_variants(int mode, int* count)
{
    if (!count) { set_error("gemm_num_variants: null out"); return CHAI_ERR_ARG; }
    *count = gemm_num_variants(mode);
    return CHAI_OK;
}

extern "C" int chai_hip_gemm_variant_params(int mode, int variant, int* params)
{
    if (!params) { set_error("gemm_variant_params: null out"); return CHAI_ERR_ARG; }
    const int rc = gemm_variant_params(mode, variant, params);
    if (rc != CHAI_OK) set_error("gemm_variant_params: no variant %d for mode %d", variant, mode);
    return rc;
}
