/*
 * chai_oracle.h -- CPU restatement of Chai's interpreter semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the MI355X
 * back end: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  Nothing in the product (chai_amd/, include/) links it,
 * calls it or falls back to it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Pinning (what independent evidence fixes each function) is listed in the
 * header of each .c file; a function with no independent pin says
 * "parity unpinned" there.
 *
 * Precision codes are Chai's PrecType codes (misc/PrecType.hpp:18-21):
 *   0 = uint32, 1 = int32, 4 = float, 8 = double.
 */
#ifndef CHAI_ORACLE_H
#define CHAI_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- counter-based RNG blocks (Random123 1.06 algorithms as emitted by
 *      Random123/R123Philox.cpp:34-286 and Random123/R123Threefry.cpp:131-290) */
void chai_or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void chai_or_philox2x32_10(const uint32_t ctr[2], const uint32_t key[1], uint32_t out[2]);
void chai_or_philox2x64_10(const uint64_t ctr[2], const uint64_t key[1], uint64_t out[2]);
/* Chai's Threefry4x32-20 (no word permutation, R123Threefry.cpp:238-255) and
 * Random123's standard form (kept to pin the shared key schedule/rotations). */
void chai_or_threefry4x32_20(const uint32_t ctr[4], const uint32_t key[4], uint32_t out[4]);
void chai_or_threefry4x32_20_r123(const uint32_t ctr[4], const uint32_t key[4], uint32_t out[4]);
void chai_or_threefry2x32_20(const uint32_t ctr[2], const uint32_t key[2], uint32_t out[2]);
void chai_or_threefry2x64_20(const uint64_t ctr[2], const uint64_t key[2], uint64_t out[2]);
/* raw words of one stream vector (key / counter layout of the stream) */
int chai_or_rng_words(int variant, int pr, int vl, uint64_t vidx, uint64_t seed, uint64_t* out);

/* ---- canonical Box-Muller math (double, no contraction, explicit fma) */
double chai_or_log(double x);
float chai_or_logf(float x);
void chai_or_sincos2pif(float u, float* s, float* c);
void   chai_or_sincos2pi(double u, double* s, double* c);

/* ---- canonical RNG stream (api/RngFun.cpp:15-96, kernel/Function.cpp:433-503,
 *      Random123/R123Philox.cpp:288-520, enqueue/EnqueueAST.cpp:496-546).
 *      out is dense [slots][H][W] of the precision's C type. */
int chai_or_rng_fill(int variant, int prec, int vec_len, uint64_t seed,
                     size_t W, size_t H, size_t slots, int normal,
                     int has_limits, double lo, double hi, void* out);

/* ---- GEMM (interp/InterpMatmulBase.hpp:49-76, interp/InterpMatmulBase.cpp:65-886).
 *      Row-major.  op(A) is M x K, op(B) is K x N.  trans_* select the
 *      transposed operand layout (jitstmt/MakeStmt.cpp:704-720).
 *      Interpreter semantics: each product is formed in the C-promoted type of
 *      the two operands, accumulated in double, rounded once to prec_c. */
int chai_or_matmul_interp(int prec_a, int prec_b, int prec_c, int trans_a, int trans_b,
                          size_t M, size_t N, size_t K,
                          const void* A, size_t lda, const void* B, size_t ldb,
                          void* C, size_t ldc, int general, double alpha, double beta);

/* FP64 reference GEMM: products formed exactly in double (exact for
 * f32 x f32 and small-int x f32), accumulated in double, C stored as double.
 * Multi-threaded with OpenMP when nthreads > 1. */
int chai_or_gemm_f64ref(int prec_a, int prec_b, int trans_a, int trans_b,
                        size_t M, size_t N, size_t K,
                        const void* A, size_t lda, const void* B, size_t ldb,
                        double* C, size_t ldc, int nthreads);

/* GEMV y = op(A) x (interp/InterpMatmulBase.hpp:38-47), interpreter semantics. */
int chai_or_matvec_interp(int prec_a, int prec_x, int prec_y, int trans_a,
                          size_t M, size_t N, const void* A, size_t lda,
                          const void* x, void* y, int general, double alpha, double beta);

/* ---- reductions (interp/InterpAccum.cpp:13-70): op 132 = sum, 131 = mean.
 *      in is [slots][n]; out is [slots] of prec's type. */
int chai_or_reduce(int op, int prec, const void* in, size_t n, size_t slots, void* out);
/* dot product (interp/InterpDotprod.cpp:14-166) accumulated in double. */
int chai_or_dot(int prec_a, int prec_b, const void* a, const void* b, size_t n,
                size_t slots, void* out);

/* ---- trace hashing (misc/TEA.cpp:13-89) */
uint64_t chai_or_tea_hash(const uint64_t* data, size_t n);

/* ---- trace evaluation with interpreter semantics (trace.c).  bufs[v] is a
 *      host buffer for every variable (inputs filled, others written). */
struct chai_trace;
int chai_or_trace_eval(const struct chai_trace* trace, void* const* bufs);

/* ---- tile-cost MLP (3 x 64 ReLU, scalar output), evaluated in double from f32
 * parameters laid out as W1 [F][64], b1, W2 [64][64], b2, W3, b3, w4 [64], b4 */
size_t chai_or_mlp_num_params(int F);
int chai_or_mlp_forward(const float* params, int F, const float* X, size_t B, double* out);
double chai_or_mlp_gradient(const float* params, int F, const float* X, const float* Y, size_t B, double* grad);

#ifdef __cplusplus
}
#endif
#endif
