/*
 * linalg.c -- oracle GEMM / GEMV / reductions / TEA hash.  TEST INFRASTRUCTURE.
 *
 * Restates:
 *   matmat  interp/InterpMatmulBase.hpp:49-61  C[y][x] = sum_z A[y][z]*B[z][x],
 *           the product formed in the C-promoted type of the two element
 *           types, the sum in a double accumulator, one rounding to the output.
 *   addgen  interp/InterpMatmulBase.hpp:63-76 and .cpp:60-62 (MATMAT_GENERAL):
 *           C = alpha*tmpAB + beta*C computed in double, tmpAB held in the
 *           promoted product type.
 *   matvec  interp/InterpMatmulBase.hpp:38-47.
 *   sum/mean interp/InterpAccum.cpp:13-70 (double accumulator, mean = sum/LEN,
 *           converted to the element type).
 *   dot     interp/InterpDotprod.cpp:14-166 (double accumulator).
 *   TEA     misc/TEA.cpp:13-31,62-89 (Wagner's TEA variant, Davies-Meyer).
 *
 * Pinned by: numpy float64 matmul of the same inputs (tests/test_oracle_linalg.py),
 * the all-ones sentinel (C == K, autotune/EvergreenMatmulMM.cpp:800-822),
 * integer closed forms, and -- for TEA -- the reference's own misc/TEA.cpp
 * compiled into oracle/_ref (tests/test_oracle_ref.py).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include "chai_oracle.h"


static double load(int prec, const void* p, size_t i)
{
    switch (prec) {
        case 0: return (double)((const uint32_t*)p)[i];
        case 1: return (double)((const int32_t*)p)[i];
        case 4: return (double)((const float*)p)[i];
        default: return ((const double*)p)[i];
    }
}

/* product in the C-promoted type of (pa, pb), returned as double */
static double promoted_product(int pa, int pb, const void* A, size_t ia, const void* B, size_t ib)
{
    int pp = pa > pb ? pa : pb;
    if (pp == 8) return load(pa, A, ia) * load(pb, B, ib);
    if (pp == 4) {
        float a = (float)load(pa, A, ia), b = (float)load(pb, B, ib);
        return (double)(a * b);
    }
    if (pa == 0 || pb == 0) {     /* unsigned arithmetic (uint x uint, uint x int) */
        /* C's usual conversions make uint32 * int32 an unsigned product, which
         * the double accumulator adds as a nonnegative value even when the
         * result type is int32 (matmat<uint32_t, int32_t, int32_t>) */
        uint32_t a = pa == 0 ? ((const uint32_t*)A)[ia] : (uint32_t)((const int32_t*)A)[ia];
        uint32_t b = pb == 0 ? ((const uint32_t*)B)[ib] : (uint32_t)((const int32_t*)B)[ib];
        return (double)(uint32_t)(a * b);
    }
    int32_t a = ((const int32_t*)A)[ia], b = ((const int32_t*)B)[ib];
    return (double)(int32_t)((uint32_t)a * (uint32_t)b);
}

/* double -> int32 / uint32 as the reference's x86-64 build converts them
 * (g++ emits cvttsd2si): to int32 a NaN or a value outside [-2^31, 2^31)
 * gives INT_MIN; to uint32 the conversion goes through int64 (cvttsd2si with
 * a 64-bit destination) and keeps the low 32 bits, so a NaN or a value
 * outside the int64 range gives 0. */
static int32_t x86_to_i32(double v)
{
    return (v != v || v >= 2147483648.0 || v < -2147483648.0) ? INT32_MIN : (int32_t)v;
}

static uint32_t x86_to_u32(double v)
{
    return (v != v || v >= 9223372036854775808.0 || v < -9223372036854775808.0)
               ? 0u : (uint32_t)(uint64_t)(int64_t)v;
}

static double round_to(int prec, double v)
{
    switch (prec) {
        case 0: return (double)x86_to_u32(v);
        case 1: return (double)x86_to_i32(v);
        case 4: return (double)(float)v;
        default: return v;
    }
}

static void store(int prec, void* p, size_t i, double v)
{
    switch (prec) {
        case 0: ((uint32_t*)p)[i] = x86_to_u32(v); break;
        case 1: ((int32_t*)p)[i] = x86_to_i32(v); break;
        case 4: ((float*)p)[i] = (float)v; break;
        default: ((double*)p)[i] = v; break;
    }
}

static int valid(int p) { return p == 0 || p == 1 || p == 4 || p == 8; }

int chai_or_matmul_interp(int pa, int pb, int pc, int ta, int tb,
                          size_t M, size_t N, size_t K,
                          const void* A, size_t lda, const void* B, size_t ldb,
                          void* C, size_t ldc, int general, double alpha, double beta)
{
    if (!valid(pa) || !valid(pb) || !valid(pc)) return -1;
    int pp = pa > pb ? pa : pb;
    for (size_t y = 0; y < M; y++)
    for (size_t x = 0; x < N; x++) {
        double acc = 0.0;
        for (size_t z = 0; z < K; z++) {
            size_t ia = ta ? z * lda + y : y * lda + z;
            size_t ib = tb ? x * ldb + z : z * ldb + x;
            acc += promoted_product(pa, pb, A, ia, B, ib);
        }
        if (general) {
            double tmp = round_to(pp, acc);
            double c = load(pc, C, y * ldc + x);
            store(pc, C, y * ldc + x, alpha * tmp + beta * c);
        } else {
            store(pc, C, y * ldc + x, acc);
        }
    }
    return 0;
}

int chai_or_gemm_f64ref(int pa, int pb, int ta, int tb, size_t M, size_t N, size_t K,
                        const void* A, size_t lda, const void* B, size_t ldb,
                        double* C, size_t ldc, int nthreads)
{
    if (!valid(pa) || !valid(pb)) return -1;
    /* blocked over z so B rows stream; each (y, x) accumulates in z order */
    const size_t NB = 256;
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
    for (size_t y = 0; y < M; y++) {
        double* crow = C + y * ldc;
        for (size_t x = 0; x < N; x++) crow[x] = 0.0;
        for (size_t x0 = 0; x0 < N; x0 += NB) {
            size_t x1 = x0 + NB < N ? x0 + NB : N;
            for (size_t z = 0; z < K; z++) {
                double a = load(pa, A, ta ? z * lda + y : y * lda + z);
                if (tb) {
                    for (size_t x = x0; x < x1; x++) crow[x] += a * load(pb, B, x * ldb + z);
                } else if (pb == 4) {
                    const float* brow = (const float*)B + z * ldb;
                    for (size_t x = x0; x < x1; x++) crow[x] += a * (double)brow[x];
                } else if (pb == 8) {
                    const double* brow = (const double*)B + z * ldb;
                    for (size_t x = x0; x < x1; x++) crow[x] += a * brow[x];
                } else {
                    for (size_t x = x0; x < x1; x++) crow[x] += a * load(pb, B, z * ldb + x);
                }
            }
        }
    }
    return 0;
}

int chai_or_matvec_interp(int pa, int px, int py, int ta, size_t M, size_t N,
                          const void* A, size_t lda, const void* x, void* y,
                          int general, double alpha, double beta)
{
    if (!valid(pa) || !valid(px) || !valid(py)) return -1;
    int pp = pa > px ? pa : px;
    for (size_t r = 0; r < M; r++) {
        double acc = 0.0;
        for (size_t c = 0; c < N; c++) {
            size_t ia = ta ? c * lda + r : r * lda + c;
            acc += promoted_product(pa, px, A, ia, x, c);
        }
        if (general) {
            double tmp = round_to(pp, acc);
            store(py, y, r, alpha * tmp + beta * load(py, y, r));
        } else {
            store(py, y, r, acc);
        }
    }
    return 0;
}

int chai_or_reduce(int op, int prec, const void* in, size_t n, size_t slots, void* out)
{
    if (!valid(prec) || (op != 131 && op != 132)) return -1;
    for (size_t s = 0; s < slots; s++) {
        double acc = 0.0;
        for (size_t j = 0; j < n; j++) acc += load(prec, in, s * n + j);
        if (op == 131) acc /= (double)n;
        store(prec, out, s, acc);
    }
    return 0;
}

int chai_or_dot(int pa, int pb, const void* a, const void* b, size_t n, size_t slots, void* out)
{
    if (!valid(pa) || !valid(pb)) return -1;
    int pp = pa > pb ? pa : pb;
    for (size_t s = 0; s < slots; s++) {
        double acc = 0.0;
        for (size_t j = 0; j < n; j++)
            acc += promoted_product(pa, pb, a, s * n + j, b, s * n + j);
        store(pp, out, s, acc);
    }
    return 0;
}

/* ------------------------------------------------------------- TEA hash */

static void tea_encipher(const uint32_t v[2], uint32_t w[2], const uint32_t k[4])
{
    uint32_t y = v[0], z = v[1], sum = 0, delta = 0x9E3779B9u;
    for (int n = 0; n < 32; n++) {
        y += (((z << 4) ^ (z >> 5)) + z) ^ (sum + k[sum & 3]);
        sum += delta;
        z += (((y << 4) ^ (y >> 5)) + y) ^ (sum + k[(sum >> 11) & 3]);
    }
    w[0] = y; w[1] = z;
}

uint64_t chai_or_tea_hash(const uint64_t* data, size_t n)
{
    uint32_t v[2] = {314159u, 271828u};
    uint32_t k[4] = {828172u, 951413u, 0u, 0u};
    uint32_t w[2] = {0u, 0u};
    for (size_t i = 0; i < n; i++) {
        uint32_t lohi[2];
        memcpy(lohi, &data[i], 8);
        k[1] = v[0];
        k[0] = v[1];
        k[3] = lohi[0];
        k[2] = lohi[1];
        tea_encipher(v, w, k);
        v[0] = w[0];
        v[1] = w[1];
    }
    uint64_t r;
    memcpy(&r, w, 8);
    return r;
}
