// gemv.hip -- chai_hip_gemv: y = op(A) x [ * alpha + beta y ].
//
// Replaces the Evergreen "MV" OpenCL generator (autotune/EvergreenMatmulMV.cpp:
// 884-1403) and the vector-matrix swap of enqueue/EnqueueMatmulAuto.cpp:282-401.
// HBM-bound: A is streamed exactly once.
//   y = A x   : one wavefront per row; each lane streams 16-byte chunks of the
//               row with non-temporal loads (4 in flight), multiplies by the
//               matching chunk of x (L2 resident) and accumulates in FP32;
//               the 64 lane partials are combined in FP64 by a butterfly.
//   y = A^T x : each thread owns 4 consecutive columns (one 16-byte load per
//               row), blocks split the rows; per-chunk FP64 partials are
//               summed in a fixed order by a second pass.
#include "common.hpp"

#include <cstring>
#include "reduce_device.hpp"

#include <type_traits>

namespace chai_hip {

typedef float gv_f4 __attribute__((ext_vector_type(4)));

template <typename T>
__device__ __forceinline__ double to_d(T v) { return (double)v; }

template <typename TA, typename TX, typename TACC>
__global__ __launch_bounds__(256) void gemv_rows(const TA* __restrict__ A, uint64_t lda,
                                                 const TX* __restrict__ x, void* __restrict__ y,
                                                 int prec_y, uint64_t M, uint64_t N, int general,
                                                 double alpha, double beta, uint64_t sa,
                                                 uint64_t sx, uint64_t sy)
{
    const uint64_t b = blockIdx.z;
    A += b * sa;
    x += b * sx;
    const int lane = threadIdx.x & 63;
    const uint64_t row = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= M) return;
    const TA* a = A + row * lda;
    TACC acc0 = 0, acc1 = 0;
    uint64_t k = 0;
    if constexpr (std::is_same<TA, float>::value && std::is_same<TX, float>::value) {
        const bool vec = ((uintptr_t)a % 16 == 0) && ((uintptr_t)x % 16 == 0);
        if (vec) {
            const gv_f4* av = reinterpret_cast<const gv_f4*>(a);
            const gv_f4* xv = reinterpret_cast<const gv_f4*>(x);
            const uint64_t nv = N / 4;
            uint64_t i = lane;
            // 4 independent 16-byte loads per lane per iteration
            for (; i + 192 < nv; i += 256) {
                const gv_f4 a0 = __builtin_nontemporal_load(av + i);
                const gv_f4 a1 = __builtin_nontemporal_load(av + i + 64);
                const gv_f4 a2 = __builtin_nontemporal_load(av + i + 128);
                const gv_f4 a3 = __builtin_nontemporal_load(av + i + 192);
                const gv_f4 x0 = xv[i], x1 = xv[i + 64], x2 = xv[i + 128], x3 = xv[i + 192];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    acc0 = __builtin_fmaf(a0[j], x0[j], acc0);
                    acc1 = __builtin_fmaf(a1[j], x1[j], acc1);
                    acc0 = __builtin_fmaf(a2[j], x2[j], acc0);
                    acc1 = __builtin_fmaf(a3[j], x3[j], acc1);
                }
            }
            for (; i < nv; i += 64) {
                const gv_f4 a0 = __builtin_nontemporal_load(av + i);
                const gv_f4 x0 = xv[i];
#pragma unroll
                for (int j = 0; j < 4; j++) acc0 = __builtin_fmaf(a0[j], x0[j], acc0);
            }
            k = nv * 4;
        }
    }
    for (uint64_t i = k + lane; i < N; i += 64) {
        if constexpr (std::is_same<TACC, float>::value)
            acc0 = __builtin_fmaf((float)a[i], (float)x[i], acc0);
        else
            acc0 = __builtin_fma((double)a[i], (double)x[i], acc0);
    }
    double r = wave_sum_f64((double)acc0 + (double)acc1);
    if (lane == 0) {
        const uint64_t yi = b * sy + row;
        if (prec_y == CHAI_F64) {
            double* yd = (double*)y;
            yd[yi] = general ? alpha * r + beta * yd[yi] : r;
        } else {
            float* yf = (float*)y;
            const float t = (float)r;   // the product held in the promoted type first
            yf[yi] = general ? (float)(alpha * (double)t + beta * (double)yf[yi]) : t;
        }
    }
}

// y = A^T x : partial[chunk][col] over row chunks
template <typename TA, typename TX>
__global__ __launch_bounds__(256) void gemv_cols_partial(const TA* __restrict__ A, uint64_t lda,
                                                         const TX* __restrict__ x, uint64_t M,
                                                         uint64_t N, uint64_t rows_per_chunk,
                                                         double* __restrict__ partial,
                                                         uint64_t sa, uint64_t sx)
{
    const uint64_t b = blockIdx.z;
    A += b * sa;
    x += b * sx;
    const uint64_t c0 = ((uint64_t)blockIdx.x * 256 + threadIdx.x) * 4;
    if (c0 >= N) return;
    const uint64_t r0 = (uint64_t)blockIdx.y * rows_per_chunk;
    const uint64_t r1 = r0 + rows_per_chunk < M ? r0 + rows_per_chunk : M;
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    double accd[4] = {0.0, 0.0, 0.0, 0.0};
    const int nc = (N - c0) < 4 ? (int)(N - c0) : 4;
    for (uint64_t r = r0; r < r1; r++) {
        const TA* a = A + r * lda + c0;
        const double xr = (double)x[r];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if (j < nc) {
                if constexpr (std::is_same<TA, double>::value || std::is_same<TX, double>::value)
                    accd[j] = __builtin_fma((double)a[j], xr, accd[j]);
                else
                    acc[j] = __builtin_fmaf((float)a[j], (float)xr, acc[j]);
            }
        }
    }
    double* p = partial + ((uint64_t)b * gridDim.y + blockIdx.y) * N + c0;
    for (int j = 0; j < nc; j++) p[j] = (double)acc[j] + accd[j];
}

__global__ __launch_bounds__(256) void gemv_cols_final(const double* __restrict__ partial,
                                                       uint64_t chunks, uint64_t N, void* y,
                                                       int prec_y, int general, double alpha,
                                                       double beta, uint64_t sy)
{
    const uint64_t b = blockIdx.z;
    const uint64_t c = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= N) return;
    double s = 0.0;
    for (uint64_t k = 0; k < chunks; k++) s += partial[(b * chunks + k) * N + c];
    const uint64_t yi = b * sy + c;
    if (prec_y == CHAI_F64) {
        double* yd = (double*)y;
        yd[yi] = general ? alpha * s + beta * yd[yi] : s;
    } else {
        float* yf = (float*)y;
        const float t = (float)s;
        yf[yi] = general ? (float)(alpha * (double)t + beta * (double)yf[yi]) : t;
    }
}

template <typename TA, typename TX>
static int gemv_typed(int prec_y, int trans_a, size_t M, size_t N, const void* A, size_t lda,
                      const void* x, void* y, int general, double alpha, double beta, size_t nb,
                      size_t sa, size_t sx, size_t sy, hipStream_t st)
{
    constexpr bool f64 = std::is_same<TA, double>::value || std::is_same<TX, double>::value;
    if (!trans_a) {
        dim3 grid((unsigned)((M + 3) / 4), 1, (unsigned)nb);
        gemv_rows<TA, TX, typename std::conditional<f64, double, float>::type>
            <<<grid, 256, 0, st>>>((const TA*)A, lda, (const TX*)x, y, prec_y, M, N, general, alpha,
                                   beta, sa, sx, sy);
        CHAI_LAUNCH_CHECK("gemv_rows");
        return CHAI_OK;
    }
    // op(A) = A^T: A stored M x N, y has N entries, x has M entries
    const uint64_t col_blocks = (N + 1023) / 1024;
    uint64_t chunks = (256 * 8) / (col_blocks * nb);
    if (chunks < 1) chunks = 1;
    if (chunks > M) chunks = M;
    const uint64_t rpc = (M + chunks - 1) / chunks;
    chunks = (M + rpc - 1) / rpc;
    double* partial = nullptr;
    CHAI_HIP_TRY(hipMallocAsync((void**)&partial, sizeof(double) * chunks * N * nb, st));
    gemv_cols_partial<TA, TX><<<dim3((unsigned)col_blocks, (unsigned)chunks, (unsigned)nb), 256, 0, st>>>(
        (const TA*)A, lda, (const TX*)x, M, N, rpc, partial, sa, sx);
    gemv_cols_final<<<dim3((unsigned)((N + 255) / 256), 1, (unsigned)nb), 256, 0, st>>>(
        partial, chunks, N, y, prec_y, general, alpha, beta, sy);
    hipError_t e = hipGetLastError();
    CHAI_HIP_TRY(hipFreeAsync(partial, st));
    if (e != hipSuccess) return hip_fail(e, "gemv_cols launch");
    return CHAI_OK;
}

template <typename TA>
static int gemv_x(int prec_x, int prec_y, int trans_a, size_t M, size_t N, const void* A, size_t lda,
                  const void* x, void* y, int general, double alpha, double beta, size_t nb,
                  size_t sa, size_t sx, size_t sy, hipStream_t st)
{
    switch (prec_x) {
        case CHAI_F32: return gemv_typed<TA, float>(prec_y, trans_a, M, N, A, lda, x, y, general, alpha, beta, nb, sa, sx, sy, st);
        case CHAI_F64: return gemv_typed<TA, double>(prec_y, trans_a, M, N, A, lda, x, y, general, alpha, beta, nb, sa, sx, sy, st);
        case CHAI_I32: return gemv_typed<TA, int32_t>(prec_y, trans_a, M, N, A, lda, x, y, general, alpha, beta, nb, sa, sx, sy, st);
        default: return gemv_typed<TA, uint32_t>(prec_y, trans_a, M, N, A, lda, x, y, general, alpha, beta, nb, sa, sx, sy, st);
    }
}

}  // namespace chai_hip

using namespace chai_hip;

extern "C" int chai_hip_gemv(int prec_a, int prec_x, int prec_y, int trans_a, size_t M, size_t N,
                             const void* A, size_t lda, const void* x, void* y, int general,
                             double alpha, double beta, size_t batch, int same_a,
                             size_t stride_a, size_t stride_x, size_t stride_y, void* stream)
{
    if (!valid_prec(prec_a) || !valid_prec(prec_x) || !valid_prec(prec_y)) {
        set_error("gemv: bad precision");
        return CHAI_ERR_ARG;
    }
    if (prec_a <= CHAI_I32 && prec_x <= CHAI_I32) {
        // the interpreter's integer product semantics: the GEMM family's
        // integer path with x as a one-column matrix
        if (M == 0 || N == 0) return CHAI_OK;
        if (!A || !x || !y) { set_error("gemv: null operand"); return CHAI_ERR_ARG; }
        chai_gemm_desc d;
        std::memset(&d, 0, sizeof d);
        d.prec_a = prec_a; d.prec_b = prec_x; d.prec_c = prec_y;
        d.trans_a = trans_a;
        d.M = trans_a ? N : M; d.K = trans_a ? M : N; d.N = 1;
        d.lda = lda; d.ldb = 1; d.ldc = 1;
        d.general = general; d.alpha = alpha; d.beta = beta;
        d.batch = batch ? batch : 1;
        d.same_a = same_a;
        d.stride_a = stride_a; d.stride_b = stride_x; d.stride_c = stride_y;
        d.mode = CHAI_GEMM_INT;
        d.variant = -1;
        return chai_hip_gemm(&d, A, x, y, stream);
    }
    const int want_y = (prec_a == CHAI_F64 || prec_x == CHAI_F64) ? CHAI_F64 : CHAI_F32;
    if (prec_y != want_y) { set_error("gemv: y must have the promoted precision (%d)", want_y); return CHAI_ERR_ARG; }
    if (M == 0 || N == 0) return CHAI_OK;
    if (!A || !x || !y) { set_error("gemv: null operand"); return CHAI_ERR_ARG; }
    if (lda < N) { set_error("gemv: lda < N"); return CHAI_ERR_ARG; }
    const size_t nb = batch ? batch : 1;
    if (nb > 65535) { set_error("gemv: at most 65535 instances"); return CHAI_ERR_ARG; }
    const size_t sa = same_a ? 0 : stride_a;
    hipStream_t st = as_stream(stream);
    switch (prec_a) {
        case CHAI_F32: return gemv_x<float>(prec_x, prec_y, trans_a, M, N, A, lda, x, y, general, alpha, beta, nb, sa, stride_x, stride_y, st);
        case CHAI_F64: return gemv_x<double>(prec_x, prec_y, trans_a, M, N, A, lda, x, y, general, alpha, beta, nb, sa, stride_x, stride_y, st);
        case CHAI_I32: return gemv_x<int32_t>(prec_x, prec_y, trans_a, M, N, A, lda, x, y, general, alpha, beta, nb, sa, stride_x, stride_y, st);
        default: return gemv_x<uint32_t>(prec_x, prec_y, trans_a, M, N, A, lda, x, y, general, alpha, beta, nb, sa, stride_x, stride_y, st);
    }
}
