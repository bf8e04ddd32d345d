// gemv.hip -- the matrix-vector family: y = op(A) x [ * alpha + beta y ].
//
// Replaces the Evergreen "MV" OpenCL generator (autotune/EvergreenMatmulMV.cpp:
// 884-1403) and the vector-matrix swap of enqueue/EnqueueMatmulAuto.cpp:282-401.
// HBM-bound: A is streamed exactly once.  Each product is formed in the
// operands' promoted type (f32 unless one is f64) and summed in FP64, as the
// interpreter's matvec does (interp/InterpMatmulBase.hpp:37-46: a double
// accumulator adds d0 * d1); the sum is rounded once to y's type.
//   y = A x   : SPLIT waves per row (4 / SPLIT rows per 256-thread block);
//               each lane streams 16-byte chunks of its share of the row
//               with non-temporal loads, LIF in flight, multiplies by the
//               matching chunk of x (L2 resident); the lane sums combine by
//               a butterfly, the SPLIT wave sums in wave order.
//   y = A^T x : each thread owns 4 consecutive columns (one 16-byte load per
//               row, LIF rows in flight), blocks split the rows into chunks
//               sized so that about 1024 x SPLIT blocks run; the per-chunk
//               FP64 partials are summed in chunk order by a second pass.
// The tile variants (SPLIT, LIF) are the tuner's search space for mode
// CHAI_GEMM_GEMV; they differ only in how the FP64 sum is ordered.
#include "common.hpp"
#include "gemm_internal.hpp"
#include "reduce_device.hpp"

#include <cstring>
#include <type_traits>

namespace chai_hip {

typedef float gv_f4 __attribute__((ext_vector_type(4)));

// the product as C++ forms it for these operand types, widened for the sum
template <typename TA, typename TX>
__device__ __forceinline__ double gv_prod(TA a, TX x)
{
    using P = typename std::conditional<std::is_same<TA, double>::value || std::is_same<TX, double>::value, double,
                                        float>::type;
    return (double)((P)a * (P)x);
}

__device__ __forceinline__ void gv_store(void* y, uint64_t yi, int prec_y, double s, int general, double alpha,
                                         double beta)
{
    if (prec_y == CHAI_F64) {
        double* yd = (double*)y;
        yd[yi] = general ? alpha * s + beta * yd[yi] : s;
    } else {
        float* yf = (float*)y;
        const float t = (float)s;   // the product held in the promoted type first
        yf[yi] = general ? (float)(alpha * (double)t + beta * (double)yf[yi]) : t;
    }
}

template <typename TA, typename TX, int SPLIT, int LIF>
__global__ __launch_bounds__(256) void gemv_rows(const TA* __restrict__ A, uint64_t lda,
                                                 const TX* __restrict__ x, void* __restrict__ y,
                                                 int prec_y, uint64_t M, uint64_t N, int general,
                                                 double alpha, double beta, uint64_t sa,
                                                 uint64_t sx, uint64_t sy)
{
    constexpr int RPB = 4 / SPLIT;                 // rows per block
    __shared__ double part[4];
    const uint64_t b = blockIdx.z;
    A += b * sa;
    x += b * sx;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t row = (uint64_t)blockIdx.x * RPB + wave / SPLIT;
    const int piece = wave % SPLIT;                // this wave's share of the row
    const bool live = row < M;
    double acc0 = 0.0, acc1 = 0.0;
    if (live) {
        const TA* a = A + row * lda;
        uint64_t k = 0;
        if constexpr (std::is_same<TA, float>::value && std::is_same<TX, float>::value) {
            if (((uintptr_t)a % 16 == 0) && ((uintptr_t)x % 16 == 0)) {
                const gv_f4* av = reinterpret_cast<const gv_f4*>(a);
                const gv_f4* xv = reinterpret_cast<const gv_f4*>(x);
                const uint64_t nv = N / 4, S = 64 * SPLIT;
                uint64_t i = (uint64_t)piece * 64 + lane;
                for (; i + (LIF - 1) * S < nv; i += LIF * S) {
                    gv_f4 ar[LIF], xr[LIF];
#pragma unroll
                    for (int u = 0; u < LIF; u++) {
                        ar[u] = __builtin_nontemporal_load(av + i + u * S);
                        xr[u] = xv[i + u * S];
                    }
#pragma unroll
                    for (int u = 0; u < LIF; u++)
#pragma unroll
                        for (int j = 0; j < 4; j++) {
                            const double p = (double)(ar[u][j] * xr[u][j]);
                            if (u & 1) acc1 += p;
                            else acc0 += p;
                        }
                }
                for (; i < nv; i += S) {
                    const gv_f4 a0 = __builtin_nontemporal_load(av + i);
                    const gv_f4 x0 = xv[i];
#pragma unroll
                    for (int j = 0; j < 4; j++) acc0 += (double)(a0[j] * x0[j]);
                }
                k = nv * 4;
            }
        }
        for (uint64_t i = k + (uint64_t)piece * 64 + lane; i < N; i += 64 * SPLIT) acc0 += gv_prod(a[i], x[i]);
    }
    double r = wave_sum_f64(acc0 + acc1);
    if constexpr (SPLIT > 1) {
        if (lane == 0) part[wave] = r;
        __syncthreads();
        if (piece != 0) return;
#pragma unroll
        for (int p = 1; p < SPLIT; p++) r += part[wave + p];
    }
    if (lane == 0 && live) gv_store(y, b * sy + row, prec_y, r, general, alpha, beta);
}

// y = A^T x : partial[chunk][col] over row chunks
template <typename TA, typename TX, int LIF>
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
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    const int nc = (N - c0) < 4 ? (int)(N - c0) : 4;
    uint64_t r = r0;
    if constexpr (std::is_same<TA, float>::value) {
        if (nc == 4 && lda % 4 == 0 && ((uintptr_t)A % 16 == 0)) {
            for (; r + LIF <= r1; r += LIF) {
                gv_f4 v[LIF];
                TX xr[LIF];
#pragma unroll
                for (int u = 0; u < LIF; u++) {
                    v[u] = __builtin_nontemporal_load(reinterpret_cast<const gv_f4*>(A + (r + u) * lda + c0));
                    xr[u] = x[r + u];
                }
#pragma unroll
                for (int u = 0; u < LIF; u++)
#pragma unroll
                    for (int j = 0; j < 4; j++) acc[j] += gv_prod(v[u][j], xr[u]);
            }
        }
    }
    for (; r < r1; r++) {
        const TA* a = A + r * lda + c0;
        const TX xr = x[r];
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (j < nc) acc[j] += gv_prod(a[j], xr);
    }
    double* p = partial + ((uint64_t)b * gridDim.y + blockIdx.y) * N + c0;
    for (int j = 0; j < nc; j++) p[j] = acc[j];
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
    gv_store(y, b * sy + c, prec_y, s, general, alpha, beta);
}

namespace {

struct GemvVariant {
    int split, lif;
};
// the tuner's search space for mode CHAI_GEMM_GEMV, in variant-id order
const GemvVariant kGemvVariants[] = {{1, 2}, {1, 4}, {1, 8}, {2, 2}, {2, 4}, {2, 8}, {4, 2}, {4, 4}, {4, 8}};
constexpr int kNumGemv = sizeof(kGemvVariants) / sizeof(kGemvVariants[0]);

struct GemvCall {
    int prec_y, trans_a, general;
    uint64_t M, N, lda, nb, sa, sx, sy;   // A stored M x N (row stride lda)
    double alpha, beta;
};

template <typename TA, typename TX, int SPLIT, int LIF>
int launch_rows(const GemvCall& c, const void* A, const void* x, void* y, hipStream_t st)
{
    constexpr int RPB = 4 / SPLIT;
    dim3 grid((unsigned)((c.M + RPB - 1) / RPB), 1, (unsigned)c.nb);
    gemv_rows<TA, TX, SPLIT, LIF><<<grid, 256, 0, st>>>((const TA*)A, c.lda, (const TX*)x, y, c.prec_y, c.M, c.N,
                                                        c.general, c.alpha, c.beta, c.sa, c.sx, c.sy);
    CHAI_LAUNCH_CHECK("gemv_rows");
    return CHAI_OK;
}

template <typename TA, typename TX, int LIF>
int launch_cols(const GemvCall& c, int split, const void* A, const void* x, void* y, hipStream_t st)
{
    // op(A) = A^T: A stored M x N, y has N entries, x has M entries
    const uint64_t col_blocks = (c.N + 1023) / 1024;
    uint64_t chunks = (1024 * (uint64_t)split) / (col_blocks * c.nb);
    if (chunks < 1) chunks = 1;
    if (chunks > c.M) chunks = c.M;
    const uint64_t rpc = (c.M + chunks - 1) / chunks;
    chunks = (c.M + rpc - 1) / rpc;
    if (col_blocks > 0x7fffffffull) { set_error("gemv: too many columns"); return CHAI_ERR_ARG; }
    double* partial = nullptr;
    CHAI_HIP_TRY(hipMallocAsync((void**)&partial, sizeof(double) * chunks * c.N * c.nb, st));
    gemv_cols_partial<TA, TX, LIF><<<dim3((unsigned)col_blocks, (unsigned)chunks, (unsigned)c.nb), 256, 0, st>>>(
        (const TA*)A, c.lda, (const TX*)x, c.M, c.N, rpc, partial, c.sa, c.sx);
    gemv_cols_final<<<dim3((unsigned)((c.N + 255) / 256), 1, (unsigned)c.nb), 256, 0, st>>>(
        partial, chunks, c.N, y, c.prec_y, c.general, c.alpha, c.beta, c.sy);
    hipError_t e = hipGetLastError();
    CHAI_HIP_TRY(hipFreeAsync(partial, st));
    if (e != hipSuccess) return hip_fail(e, "gemv_cols launch");
    return CHAI_OK;
}

template <typename TA, typename TX, int SPLIT>
int launch_split(const GemvCall& c, int lif, const void* A, const void* x, void* y, hipStream_t st)
{
    if (c.trans_a) {
        switch (lif) {
            case 2: return launch_cols<TA, TX, 2>(c, SPLIT, A, x, y, st);
            case 4: return launch_cols<TA, TX, 4>(c, SPLIT, A, x, y, st);
            default: return launch_cols<TA, TX, 8>(c, SPLIT, A, x, y, st);
        }
    }
    switch (lif) {
        case 2: return launch_rows<TA, TX, SPLIT, 2>(c, A, x, y, st);
        case 4: return launch_rows<TA, TX, SPLIT, 4>(c, A, x, y, st);
        default: return launch_rows<TA, TX, SPLIT, 8>(c, A, x, y, st);
    }
}

template <typename TA, typename TX>
int launch_typed(const GemvCall& c, const GemvVariant& v, const void* A, const void* x, void* y, hipStream_t st)
{
    switch (v.split) {
        case 1: return launch_split<TA, TX, 1>(c, v.lif, A, x, y, st);
        case 2: return launch_split<TA, TX, 2>(c, v.lif, A, x, y, st);
        default: return launch_split<TA, TX, 4>(c, v.lif, A, x, y, st);
    }
}

template <typename TA>
int launch_x(int prec_x, const GemvCall& c, const GemvVariant& v, const void* A, const void* x, void* y,
             hipStream_t st)
{
    switch (prec_x) {
        case CHAI_F32: return launch_typed<TA, float>(c, v, A, x, y, st);
        case CHAI_F64: return launch_typed<TA, double>(c, v, A, x, y, st);
        case CHAI_I32: return launch_typed<TA, int32_t>(c, v, A, x, y, st);
        default: return launch_typed<TA, uint32_t>(c, v, A, x, y, st);
    }
}

}  // namespace

int gemv_num_variants() { return kNumGemv; }

int gemv_variant_params(int vid, int* p)
{
    if (vid < 0 || vid >= kNumGemv) return CHAI_ERR_ARG;
    p[0] = kGemvVariants[vid].split;
    p[1] = kGemvVariants[vid].lif;
    p[2] = p[3] = p[4] = 0;
    return CHAI_OK;
}

// Without tuning: one wave per row with 4 loads in flight when the rows fill
// the chip (at least 4096 of them over the batch), else each row split over
// 4 waves; for A^T, about 2048 blocks with 4 rows in flight.
int gemv_default_variant(const chai_gemm_desc& d)
{
    const uint64_t nb = d.batch ? d.batch : 1;
    if (d.trans_a) return 4;                          // (2, 4)
    return d.M * nb >= 4096 ? 1 : 7;                  // (1, 4) : (4, 4)
}

// The GEMV family under a GEMM descriptor whose op(B) is one contiguous
// column and whose C is one contiguous column (checked by the caller): A is
// op(A) M x K, x the K entries of B, y the M entries of C.
int gemv_launch(const chai_gemm_desc& d, const void* A, const void* x, void* y, int variant, hipStream_t st)
{
    if (variant < 0 || variant >= kNumGemv) { set_error("gemv: variant %d out of range", variant); return CHAI_ERR_ARG; }
    const uint64_t nb = d.batch ? d.batch : 1;
    if (nb > 65535) { set_error("gemv: at most 65535 instances"); return CHAI_ERR_ARG; }
    GemvCall c;
    c.prec_y = d.prec_c;
    c.trans_a = d.trans_a;
    c.general = d.general;
    c.alpha = d.alpha;
    c.beta = d.beta;
    c.M = d.trans_a ? d.K : d.M;        // rows of the stored A
    c.N = d.trans_a ? d.M : d.K;
    c.lda = d.lda;
    c.nb = nb;
    c.sa = d.same_a ? 0 : d.stride_a;
    c.sx = d.same_b ? 0 : d.stride_b;
    c.sy = d.stride_c;
    if (!d.trans_a && ceil_div_u(c.M, 4) > 0x7fffffffull) { set_error("gemv: too many rows"); return CHAI_ERR_ARG; }
    const GemvVariant& v = kGemvVariants[variant];
    switch (d.prec_a) {
        case CHAI_F32: return launch_x<float>(d.prec_b, c, v, A, x, y, st);
        case CHAI_F64: return launch_x<double>(d.prec_b, c, v, A, x, y, st);
        case CHAI_I32: return launch_x<int32_t>(d.prec_b, c, v, A, x, y, st);
        default: return launch_x<uint32_t>(d.prec_b, c, v, A, x, y, st);
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
    if (M == 0 || N == 0) return CHAI_OK;
    if (!A || !x || !y) { set_error("gemv: null operand"); return CHAI_ERR_ARG; }
    // op(A) x as a GEMM with one column: A stored M x N
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
    // integer x integer: the interpreter's integer product semantics on the
    // GEMM family's integer path; otherwise the matrix-vector kernels
    d.mode = (prec_a <= CHAI_I32 && prec_x <= CHAI_I32) ? CHAI_GEMM_INT : CHAI_GEMM_GEMV;
    d.variant = -1;
    return chai_hip_gemm(&d, A, x, y, stream);
}
