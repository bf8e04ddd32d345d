// igemm.hip -- integer x integer GEMM with the interpreter's semantics.
//
// The reference's interpreter multiplies u32 / i32 arrays as C++ does
// (interp/InterpMatmulBase.hpp:49-61): each product is formed in the operand
// types' C-promoted type -- uint32 when either operand is uint32, else int32 --
// and wraps modulo 2^32; the products are summed in a double accumulator in k
// order; the sum is converted once to the result type as the reference's x86
// build converts (cvttsd2si: to int32 a value outside [-2^31, 2^31) gives
// INT_MIN; to uint32 the conversion goes through int64 and keeps the low 32
// bits).  matmulG forms tmp = that conversion to the promoted type, then
// alpha tmp + beta C in double (InterpMatmulBase.hpp:63-76), converted to C's
// type.  Its OpenCL kernels instead accumulate in the promoted 32-bit type
// (autotune/EvergreenMatmulMM.cpp:1159); the interpreter is the contract here.
//
// Exactness: the double accumulator is exact while every partial sum stays
// below 2^53, i.e. for K <= 2^21 (products below 2^32 in magnitude).  There the
// kernel accumulates the products in int64 (same value, integer adds);
// beyond, it accumulates in double in k order, rounding as the interpreter
// does.  Integer products are VALU work (v_mul_lo_u32, 64-bit adds): this path
// serves parity of integer programs, not throughput.
#include "common.hpp"
#include "gemm_internal.hpp"

#include <type_traits>

namespace chai_hip {

namespace {

constexpr int IG_T = 64;      // output tile edge
constexpr int IG_K = 32;      // k chunk staged in LDS

__device__ __forceinline__ int32_t x86_i32(double v)
{
    return (v != v || v >= 0x1p31 || v < -0x1p31) ? INT32_MIN : (int32_t)v;
}

__device__ __forceinline__ uint32_t x86_u32(double v)
{
    return (v != v || v >= 0x1p63 || v < -0x1p63) ? 0u : (uint32_t)(unsigned long long)(long long)v;
}

template <typename TC>
__device__ __forceinline__ TC to_c(double v)
{
    if constexpr (std::is_same<TC, uint32_t>::value) return x86_u32(v);
    else if constexpr (std::is_same<TC, int32_t>::value) return x86_i32(v);
    else return (TC)v;
}

template <typename TC>
__device__ __forceinline__ double from_c(TC v) { return (double)v; }

// the wrapped product in the promoted type, as a value the accumulator adds
template <bool UNS>
__device__ __forceinline__ long long wprod(uint32_t a, uint32_t b)
{
    const uint32_t p = a * b;
    return UNS ? (long long)p : (long long)(int32_t)p;
}

// 256 threads, 16 x 16, each 4 x 4 outputs of a 64 x 64 tile; k in order.
template <typename TC, bool UNS, bool DBL>
__global__ __launch_bounds__(256) void igemm_kernel(const uint32_t* __restrict__ A, const uint32_t* __restrict__ B,
                                                    TC* __restrict__ C, uint64_t M, uint64_t N, uint64_t K,
                                                    uint64_t lda, uint64_t ldb, uint64_t ldc, int ta, int tb,
                                                    uint64_t sa, uint64_t sb, uint64_t sc, int general, int prom_u32,
                                                    double alpha, double beta)
{
    __shared__ uint32_t As[IG_K][IG_T + 1];   // As[k][m]
    __shared__ uint32_t Bs[IG_K][IG_T + 1];   // Bs[k][n]
    const uint64_t z = blockIdx.z;
    A += z * sa;
    B += z * sb;
    C += z * sc;
    const uint64_t m0 = (uint64_t)blockIdx.y * IG_T, n0 = (uint64_t)blockIdx.x * IG_T;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    long long acc[4][4];
    double accd[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) { acc[i][j] = 0; accd[i][j] = 0.0; }
    for (uint64_t k0 = 0; k0 < K; k0 += IG_K) {
        // stage op(A)[m0.., k0..] and op(B)[k0.., n0..]; 2048 elements each
        for (int e = threadIdx.x; e < IG_K * IG_T; e += 256) {
            const int kk = ta ? e / IG_T : e % IG_K, mm = ta ? e % IG_T : e / IG_K;
            const uint64_t m = m0 + mm, k = k0 + kk;
            As[kk][mm] = (m < M && k < K) ? (ta ? A[k * lda + m] : A[m * lda + k]) : 0u;
            const int kb = tb ? e % IG_K : e / IG_T, nn = tb ? e / IG_K : e % IG_T;
            const uint64_t n = n0 + nn, kq = k0 + kb;
            Bs[kb][nn] = (n < N && kq < K) ? (tb ? B[n * ldb + kq] : B[kq * ldb + n]) : 0u;
        }
        __syncthreads();
        const int kn = (int)((K - k0) < (uint64_t)IG_K ? (K - k0) : (uint64_t)IG_K);
        for (int kk = 0; kk < kn; kk++) {
            uint32_t a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; i++) a[i] = As[kk][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; j++) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const long long p = wprod<UNS>(a[i], b[j]);
                    if (DBL) accd[i][j] += (double)p;
                    else acc[i][j] += p;
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint64_t m = m0 + ty * 4 + i;
        if (m >= M) continue;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint64_t n = n0 + tx * 4 + j;
            if (n >= N) continue;
            const double s = DBL ? accd[i][j] : (double)acc[i][j];
            TC* cp = C + m * ldc + n;
            if (general) {
                // tmp in the promoted type, then alpha tmp + beta C in double
                const double tmp = prom_u32 ? (double)x86_u32(s) : (double)x86_i32(s);
                *cp = to_c<TC>(alpha * tmp + beta * from_c<TC>(*cp));
            } else {
                *cp = to_c<TC>(s);
            }
        }
    }
}

template <typename TC>
int launch_typed(const chai_gemm_desc& d, const void* A, const void* B, void* C, hipStream_t st)
{
    const bool uns = d.prec_a == CHAI_U32 || d.prec_b == CHAI_U32;
    const bool dbl = d.K > (1ull << 21);
    const int prom_u32 = d.prec_a == CHAI_U32 && d.prec_b == CHAI_U32;
    const uint64_t nb = d.batch ? d.batch : 1;
    const dim3 grid(ceil_div_u(d.N, IG_T), ceil_div_u(d.M, IG_T), (unsigned)nb);
    if (grid.y > 65535) { set_error("gemm: %zu rows exceed the integer tile grid", d.M); return CHAI_ERR_ARG; }
    const uint64_t sa = d.same_a ? 0 : d.stride_a, sb = d.same_b ? 0 : d.stride_b;
    const uint32_t* a = (const uint32_t*)A;
    const uint32_t* b = (const uint32_t*)B;
#define IGEMM_GO(U, D)                                                                                   \
    igemm_kernel<TC, U, D><<<grid, 256, 0, st>>>(a, b, (TC*)C, d.M, d.N, d.K, d.lda, d.ldb, d.ldc,       \
                                                 d.trans_a, d.trans_b, sa, sb, d.stride_c, d.general,     \
                                                 prom_u32, d.alpha, d.beta)
    if (uns && dbl) IGEMM_GO(true, true);
    else if (uns) IGEMM_GO(true, false);
    else if (dbl) IGEMM_GO(false, true);
    else IGEMM_GO(false, false);
#undef IGEMM_GO
    CHAI_LAUNCH_CHECK("igemm");
    return CHAI_OK;
}

}  // namespace

int igemm_launch(const chai_gemm_desc& d, const void* A, const void* B, void* C, hipStream_t st)
{
    switch (d.prec_c) {
        case CHAI_U32: return launch_typed<uint32_t>(d, A, B, C, st);
        case CHAI_I32: return launch_typed<int32_t>(d, A, B, C, st);
        case CHAI_F32: return launch_typed<float>(d, A, B, C, st);
        default: return launch_typed<double>(d, A, B, C, st);
    }
}

}  // namespace chai_hip
