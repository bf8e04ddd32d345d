// mlp.hip -- the learned tile-cost model: a 3 x 64 FP32 MLP with hand-written
// gfx950 forward / backward / update kernels, and the GEMM feature map the
// tuner feeds it.
//
// The reference picks GEMM parameters only by search (StandardEM::searchLoop,
// autotune/StandardEM.cpp:69-121, AutoTuneEM, autotune/AutoTuneEM.cpp:48-131)
// and remembers the winner per exogenous key (enqueue/JustInTimeMemo.cpp:
// 92-114).  The model predicts the core-kernel time of every tile variant of a
// shape it has never seen, so a new shape gets its variant from one inference
// instead of a timed search; the search remains the fallback and measures the
// model's regret.
//
// Network: x (F <= 32 features) -> 64 ReLU -> 64 ReLU -> 64 ReLU -> 1, FP32.
// Kernel layout: one wavefront per sample, lane j = hidden unit j (64 units =
// 64 lanes).  Weights sit in LDS (plus transposed copies for the backward
// pass so that every LDS read is bank-conflict free); an activation of unit k
// reaches all lanes through v_readlane.  Each wave accumulates its samples'
// gradients in registers (lane j owns column j of every weight matrix) and
// writes one partial gradient; a second kernel sums the partials in a fixed
// order (deterministic) and applies SGD or Adam.
#include "common.hpp"
#include "gemm_internal.hpp"
#include "reduce_device.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

namespace chai_hip {

constexpr int MLP_H = 64;
constexpr int MLP_MAXF = 32;
constexpr int MLP_THREADS = 256;   // 4 waves per block
constexpr int MLP_MAX_WAVES = 256;

struct MlpLayout {
    uint32_t F, W1, b1, W2, b2, W3, b3, w4, b4, P;
};

static MlpLayout mlp_layout(int F)
{
    MlpLayout L;
    L.F = (uint32_t)F;
    L.W1 = 0;                         // [F][64]
    L.b1 = L.W1 + F * MLP_H;
    L.W2 = L.b1 + MLP_H;              // [64][64], row = input unit
    L.b2 = L.W2 + MLP_H * MLP_H;
    L.W3 = L.b2 + MLP_H;
    L.b3 = L.W3 + MLP_H * MLP_H;
    L.w4 = L.b3 + MLP_H;              // [64]
    L.b4 = L.w4 + MLP_H;
    L.P = L.b4 + 1;
    return L;
}

__device__ __forceinline__ float lane_bcast(float v, int k)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}

// TRAIN = false: out[s] = y(x_s).  TRAIN = true: also accumulate the gradient
// of mean_s (y_s - t_s)^2 into partial[wave][P] (and out[s] = y_s if out).
template <bool TRAIN>
__global__ __launch_bounds__(MLP_THREADS) void mlp_kernel(const float* __restrict__ params, MlpLayout L,
                                                           const float* __restrict__ X,
                                                           const float* __restrict__ Y, uint64_t B,
                                                           float* __restrict__ out, float* __restrict__ partial,
                                                           float inv_b)
{
    extern __shared__ float sm[];
    // LDS: W1 [F][64] | b1 | W2 | b2 | W3 | b3 | w4 | b4 (same layout as params) | W2T | W3T
    const int t = threadIdx.x, lane = t & 63;
    for (uint32_t i = t; i < L.P; i += MLP_THREADS) sm[i] = params[i];
    float* W2T = sm + L.P;
    float* W3T = W2T + MLP_H * MLP_H;
    if (TRAIN) {
        for (int i = t; i < MLP_H * MLP_H; i += MLP_THREADS) {
            const int k = i >> 6, j = i & 63;   // W[k][j] -> WT[j][k]
            W2T[j * MLP_H + k] = params[L.W2 + i];
            W3T[j * MLP_H + k] = params[L.W3 + i];
        }
    }
    __syncthreads();
    const float* W1 = sm + L.W1;
    const float* W2 = sm + L.W2;
    const float* W3 = sm + L.W3;
    const float b1 = sm[L.b1 + lane], b2 = sm[L.b2 + lane], b3 = sm[L.b3 + lane];
    const float w4 = sm[L.w4 + lane], b4 = sm[L.b4];
    const int F = (int)L.F;

    float gW1[MLP_MAXF], gW2[MLP_H], gW3[MLP_H];
    float gb1 = 0.f, gb2 = 0.f, gb3 = 0.f, gw4 = 0.f, gb4 = 0.f;
    if (TRAIN) {
#pragma unroll
        for (int f = 0; f < MLP_MAXF; f++) gW1[f] = 0.f;
#pragma unroll
        for (int k = 0; k < MLP_H; k++) { gW2[k] = 0.f; gW3[k] = 0.f; }
    }

    const uint64_t wave = (uint64_t)blockIdx.x * (MLP_THREADS / 64) + (t >> 6);
    const uint64_t nw = (uint64_t)gridDim.x * (MLP_THREADS / 64);
    for (uint64_t s = wave; s < B; s += nw) {
        const float xl = lane < F ? X[s * F + lane] : 0.f;
        float z1 = b1;
#pragma unroll
        for (int f = 0; f < MLP_MAXF; f++)
            if (f < F) z1 = __builtin_fmaf(lane_bcast(xl, f), W1[f * MLP_H + lane], z1);
        const float h1 = z1 > 0.f ? z1 : 0.f;
        float z2 = b2;
#pragma unroll
        for (int k = 0; k < MLP_H; k++) z2 = __builtin_fmaf(lane_bcast(h1, k), W2[k * MLP_H + lane], z2);
        const float h2 = z2 > 0.f ? z2 : 0.f;
        float z3 = b3;
#pragma unroll
        for (int k = 0; k < MLP_H; k++) z3 = __builtin_fmaf(lane_bcast(h2, k), W3[k * MLP_H + lane], z3);
        const float h3 = z3 > 0.f ? z3 : 0.f;
        const float y = wave_sum_f32(h3 * w4) + b4;
        if (out && lane == 0) out[s] = y;
        if (!TRAIN) continue;

        const float dy = 2.f * (y - Y[s]) * inv_b;
        gb4 += dy;
        gw4 = __builtin_fmaf(h3, dy, gw4);
        const float dz3 = z3 > 0.f ? w4 * dy : 0.f;
        gb3 += dz3;
#pragma unroll
        for (int k = 0; k < MLP_H; k++) gW3[k] = __builtin_fmaf(lane_bcast(h2, k), dz3, gW3[k]);
        float a2 = 0.f;   // sum_j W3[lane][j] dz3[j]
#pragma unroll
        for (int j = 0; j < MLP_H; j++) a2 = __builtin_fmaf(lane_bcast(dz3, j), W3T[j * MLP_H + lane], a2);
        const float dz2 = z2 > 0.f ? a2 : 0.f;
        gb2 += dz2;
#pragma unroll
        for (int k = 0; k < MLP_H; k++) gW2[k] = __builtin_fmaf(lane_bcast(h1, k), dz2, gW2[k]);
        float a1 = 0.f;
#pragma unroll
        for (int j = 0; j < MLP_H; j++) a1 = __builtin_fmaf(lane_bcast(dz2, j), W2T[j * MLP_H + lane], a1);
        const float dz1 = z1 > 0.f ? a1 : 0.f;
        gb1 += dz1;
#pragma unroll
        for (int f = 0; f < MLP_MAXF; f++)
            if (f < F) gW1[f] = __builtin_fmaf(lane_bcast(xl, f), dz1, gW1[f]);
    }
    if (!TRAIN) return;
    float* g = partial + wave * L.P;
#pragma unroll
    for (int f = 0; f < MLP_MAXF; f++)
        if (f < F) g[L.W1 + f * MLP_H + lane] = gW1[f];
    g[L.b1 + lane] = gb1;
#pragma unroll
    for (int k = 0; k < MLP_H; k++) {
        g[L.W2 + k * MLP_H + lane] = gW2[k];
        g[L.W3 + k * MLP_H + lane] = gW3[k];
    }
    g[L.b2 + lane] = gb2;
    g[L.b3 + lane] = gb3;
    g[L.w4 + lane] = gw4;
    if (lane == 0) g[L.b4] = gb4;
}

// grad[p] = sum of the waves' partials (fixed order); then optionally an
// SGD (kind 0) or Adam (kind 1) step on params.
__global__ __launch_bounds__(256) void mlp_reduce_update(const float* __restrict__ partial, uint32_t nw,
                                                          uint32_t P, float* __restrict__ grad_out,
                                                          float* __restrict__ params, float* __restrict__ m,
                                                          float* __restrict__ v, int kind, float lr,
                                                          float bc1, float bc2)
{
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= P) return;
    float g = 0.f;
    for (uint32_t w = 0; w < nw; w++) g += partial[(uint64_t)w * P + p];
    if (grad_out) grad_out[p] = g;
    if (kind == 0) {
        params[p] -= lr * g;
    } else if (kind == 1) {
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        const float mm = b1 * m[p] + (1.f - b1) * g;
        const float vv = b2 * v[p] + (1.f - b2) * g * g;
        m[p] = mm;
        v[p] = vv;
        params[p] -= lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
    }
}

struct Mlp {
    MlpLayout L;
    float* params = nullptr;   // device
    float* m = nullptr;        // Adam moments
    float* v = nullptr;
    float* partial = nullptr;  // [MLP_MAX_WAVES][P]
    uint64_t steps = 0;
    ~Mlp()
    {
        if (params) (void)hipFree(params);
        if (m) (void)hipFree(m);
        if (v) (void)hipFree(v);
        if (partial) (void)hipFree(partial);
    }
};

static size_t lds_bytes(const MlpLayout& L, bool train)
{
    return sizeof(float) * (L.P + (train ? 2 * MLP_H * MLP_H : 0));
}

static int mlp_run(Mlp* N, const float* X, const float* Y, size_t B, float* out, bool train, hipStream_t st,
                   uint32_t* nw_out)
{
    const uint64_t waves_needed = B;   // one sample per wave at most per pass
    uint64_t blocks = (waves_needed + 3) / 4;
    if (train) blocks = std::min<uint64_t>(blocks, MLP_MAX_WAVES / 4);
    else blocks = std::min<uint64_t>(blocks, 4096);
    if (blocks == 0) blocks = 1;
    const size_t lds = lds_bytes(N->L, train);
    if (train) {
        static bool attr = false;
        if (!attr) {
            CHAI_HIP_TRY(hipFuncSetAttribute((const void*)mlp_kernel<true>,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes(mlp_layout(MLP_MAXF), true)));
            attr = true;
        }
        mlp_kernel<true><<<(unsigned)blocks, MLP_THREADS, lds, st>>>(N->params, N->L, X, Y, B, out, N->partial,
                                                                        B ? 1.f / (float)B : 0.f);
        CHAI_LAUNCH_CHECK("mlp_kernel<train>");
    } else {
        mlp_kernel<false><<<(unsigned)blocks, MLP_THREADS, lds, st>>>(N->params, N->L, X, Y, B, out, nullptr, 0.f);
        CHAI_LAUNCH_CHECK("mlp_kernel<forward>");
    }
    if (nw_out) *nw_out = (uint32_t)(blocks * 4);
    return CHAI_OK;
}

// ------------------------------------------------------------ GEMM features
// The families the model covers, in one-hot order.
static int family_slot(int mode)
{
    switch (mode) {
        case CHAI_GEMM_BF16X3: return 0;
        case CHAI_GEMM_FMA: return 1;
        case CHAI_GEMM_F64: return 2;
        case CHAI_GEMM_GEMV: return 3;
        case CHAI_GEMM_BF16X6: return 4;
        default: return -1;
    }
}

// The model's input for (shape, family, variant), every feature O(1): the
// shape and operand types, the family, the variant's tile parameters, and a
// crude occupancy model of the launch (resident blocks per CU from LDS and
// waves, rounds of blocks over 256 CUs, padding waste).
static void gemm_features(const chai_gemm_desc& d, int mode, int variant, float* f)
{
    const int fam = family_slot(mode);
    int p[5] = {0, 0, 0, 0, 0};
    (void)gemm_variant_params(mode, variant, p);
    const double M = (double)d.M, N = (double)d.N, K = (double)d.K;
    const double nb = (double)(d.batch ? d.batch : 1);
    for (int i = 0; i < CHAI_GEMM_FEATURES; i++) f[i] = 0.f;
    f[0] = (float)(std::log2(M) / 14);
    f[1] = (float)(std::log2(N) / 14);
    f[2] = (float)(std::log2(K) / 14);
    f[3] = (float)(d.trans_a != 0);
    f[4] = (float)(d.trans_b != 0);
    f[5] = (float)(d.prec_a == CHAI_I32 || d.prec_a == CHAI_U32);
    f[6] = (float)(d.prec_b == CHAI_I32 || d.prec_b == CHAI_U32);
    f[7] = (float)(d.prec_a == CHAI_F64 || d.prec_b == CHAI_F64);
    if (fam >= 0) f[8 + fam] = 1.f;
    double BM = 1, BN = 1, BK = 1, waves = 4, lds = 1024, kq = 1, aspect = 0, tiles = 1, used = 1;
    if (fam == 3) {
        const double split = p[0], lif = p[1];
        if (!d.trans_a) {           // rows of A: 4 / split per block
            BM = 4 / split; BN = 1; BK = 256 * lif * split;
            tiles = std::ceil(M / BM) * nb;
            used = M / (std::ceil(M / BM) * BM);
        } else {                    // 1024 columns per block, row chunks for ~1024 split blocks
            const double cb = std::ceil(M / 1024);
            const double chunks = std::max(1.0, std::min(K, std::floor(1024 * split / (cb * nb))));
            BM = 1024; BN = 1; BK = std::ceil(K / chunks);
            tiles = cb * chunks * nb;
            used = M / (cb * 1024);
        }
        waves = 4; lds = 64;
        aspect = std::log2(split) / 2;
        f[18] = (float)(std::log2(lif) / 3);
    } else {
        BM = p[0] > 0 ? p[0] : 128; BN = p[1] > 0 ? p[1] : 128; BK = p[2] > 0 ? p[2] : 16;
        waves = std::max(1, p[3] * p[4]);
        if (fam == 0) { lds = 2.0 * (2 * BM + 2 * BN) * BK * 2; kq = 64; }
        else if (fam == 4) { lds = 2.0 * (3 * BM + 3 * BN) * BK * 2; kq = 64; }
        else if (fam == 1) { lds = 2.0 * 16 * (BM + BN + 8) * 4; kq = 16; waves = 4; }
        else { lds = 2.0 * 8 * (BM + BN + 32) * 8; kq = 32; }
        aspect = std::log2((BM / std::max(1, p[3])) / (BN / std::max(1, p[4]))) / 2;
        const double tm = std::ceil(M / BM), tn = std::ceil(N / BN);
        tiles = tm * tn * nb;
        used = M * N / (tm * BM * tn * BN);
    }
    double per_cu = std::floor(160.0 * 1024 / lds);
    per_cu = std::max(1.0, std::min(per_cu, std::floor(32.0 / waves)));
    const double rounds = tiles / (256.0 * per_cu);
    const double Kp = std::ceil(K / kq) * kq;
    f[13] = (float)(std::log2(BM) / 10);
    f[14] = (float)(std::log2(BN) / 8);
    f[15] = (float)(std::log2(BK) / 14);
    f[16] = (float)(waves / 8);
    f[17] = (float)aspect;
    f[19] = (float)(std::log2(tiles) / 16);
    f[20] = (float)(std::log2(rounds) / 8);
    f[21] = (float)(rounds / std::ceil(rounds));       // last-round occupancy
    f[22] = (float)used;                               // useful fraction of the tile grid
    f[23] = (float)(K / Kp);
    f[24] = (float)(std::log2(nb) / 8);
}

}  // namespace chai_hip

using namespace chai_hip;

extern "C" {

int chai_hip_mlp_create(int inputs, uint64_t seed, void** mlp)
{
    if (!mlp || inputs < 1 || inputs > MLP_MAXF) { set_error("mlp_create: 1..%d inputs", MLP_MAXF); return CHAI_ERR_ARG; }
    auto* N = new Mlp;
    N->L = mlp_layout(inputs);
    const size_t P = N->L.P;
    // He-uniform weights (ReLU), zero biases; deterministic from the seed
    std::vector<float> h(P, 0.f);
    std::mt19937_64 g(seed);
    auto fill = [&](uint32_t off, int fan_in, int count) {
        const double a = std::sqrt(6.0 / fan_in);
        std::uniform_real_distribution<double> U(-a, a);
        for (int i = 0; i < count; i++) h[off + i] = (float)U(g);
    };
    fill(N->L.W1, inputs, inputs * MLP_H);
    fill(N->L.W2, MLP_H, MLP_H * MLP_H);
    fill(N->L.W3, MLP_H, MLP_H * MLP_H);
    fill(N->L.w4, MLP_H, MLP_H);
    hipError_t e = hipMalloc((void**)&N->params, P * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&N->m, P * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&N->v, P * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&N->partial, (size_t)MLP_MAX_WAVES * P * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(N->params, h.data(), P * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(N->m, 0, P * sizeof(float));
    if (e == hipSuccess) e = hipMemset(N->v, 0, P * sizeof(float));
    if (e != hipSuccess) { delete N; return hip_fail(e, "mlp_create"); }
    *mlp = N;
    return CHAI_OK;
}

int chai_hip_mlp_num_params(void* mlp, size_t* count)
{
    if (!mlp || !count) { set_error("mlp_num_params: null argument"); return CHAI_ERR_ARG; }
    *count = ((Mlp*)mlp)->L.P;
    return CHAI_OK;
}

int chai_hip_mlp_get_params(void* mlp, float* host, size_t count)
{
    Mlp* N = (Mlp*)mlp;
    if (!N || !host || count != N->L.P) { set_error("mlp_get_params: bad arguments"); return CHAI_ERR_ARG; }
    CHAI_HIP_TRY(hipMemcpy(host, N->params, count * sizeof(float), hipMemcpyDeviceToHost));
    return CHAI_OK;
}

int chai_hip_mlp_set_params(void* mlp, const float* host, size_t count)
{
    Mlp* N = (Mlp*)mlp;
    if (!N || !host || count != N->L.P) { set_error("mlp_set_params: bad arguments"); return CHAI_ERR_ARG; }
    CHAI_HIP_TRY(hipMemcpy(N->params, host, count * sizeof(float), hipMemcpyHostToDevice));
    CHAI_HIP_TRY(hipMemset(N->m, 0, count * sizeof(float)));
    CHAI_HIP_TRY(hipMemset(N->v, 0, count * sizeof(float)));
    N->steps = 0;
    return CHAI_OK;
}

int chai_hip_mlp_forward(void* mlp, const float* X, size_t B, float* out, void* stream)
{
    Mlp* N = (Mlp*)mlp;
    if (!N || (B && (!X || !out))) { set_error("mlp_forward: bad arguments"); return CHAI_ERR_ARG; }
    if (B == 0) return CHAI_OK;
    return mlp_run(N, X, nullptr, B, out, false, as_stream(stream), nullptr);
}

int chai_hip_mlp_backward(void* mlp, const float* X, const float* Y, size_t B, float* grad, void* stream)
{
    Mlp* N = (Mlp*)mlp;
    if (!N || !X || !Y || !grad || B == 0) { set_error("mlp_backward: bad arguments"); return CHAI_ERR_ARG; }
    hipStream_t st = as_stream(stream);
    uint32_t nw = 0;
    int rc = mlp_run(N, X, Y, B, nullptr, true, st, &nw);
    if (rc != CHAI_OK) return rc;
    mlp_reduce_update<<<ceil_div_u(N->L.P, 256), 256, 0, st>>>(N->partial, nw, N->L.P, grad, N->params, N->m,
                                                                N->v, -1, 0.f, 1.f, 1.f);
    CHAI_LAUNCH_CHECK("mlp_reduce_update");
    return CHAI_OK;
}

static int mlp_step(Mlp* N, uint32_t nw, const float* grad, int kind, float lr, hipStream_t st)
{
    N->steps++;
    const float bc1 = 1.f - std::pow(0.9f, (float)N->steps), bc2 = 1.f - std::pow(0.999f, (float)N->steps);
    mlp_reduce_update<<<ceil_div_u(N->L.P, 256), 256, 0, st>>>(grad ? grad : N->partial, grad ? 1u : nw, N->L.P,
                                                                nullptr, N->params, N->m, N->v, kind, lr, bc1, bc2);
    CHAI_LAUNCH_CHECK("mlp_reduce_update");
    return CHAI_OK;
}

int chai_hip_mlp_update(void* mlp, const float* grad, int kind, float lr, void* stream)
{
    Mlp* N = (Mlp*)mlp;
    if (!N || !grad || (kind != 0 && kind != 1)) { set_error("mlp_update: bad arguments"); return CHAI_ERR_ARG; }
    return mlp_step(N, 1, grad, kind, lr, as_stream(stream));
}

int chai_hip_mlp_train_step(void* mlp, const float* X, const float* Y, size_t B, int kind, float lr, void* stream)
{
    Mlp* N = (Mlp*)mlp;
    if (!N || !X || !Y || B == 0 || (kind != 0 && kind != 1)) { set_error("mlp_train_step: bad arguments"); return CHAI_ERR_ARG; }
    hipStream_t st = as_stream(stream);
    uint32_t nw = 0;
    int rc = mlp_run(N, X, Y, B, nullptr, true, st, &nw);
    return rc != CHAI_OK ? rc : mlp_step(N, nw, nullptr, kind, lr, st);
}

int chai_hip_mlp_release(void* mlp)
{
    delete (Mlp*)mlp;
    return CHAI_OK;
}

// the family a descriptor's mode stands for (AUTO resolves as chai_hip_gemm does)
static int model_mode(const chai_gemm_desc& d)
{
    return d.mode != CHAI_GEMM_AUTO ? d.mode : gemm_auto_mode(d);
}

int chai_hip_gemm_features(const chai_gemm_desc* d, int variant, float* out)
{
    if (!d || !out || !d->M || !d->N || !d->K) { set_error("gemm_features: bad arguments"); return CHAI_ERR_ARG; }
    const int mode = model_mode(*d);
    if (family_slot(mode) < 0 || variant < 0 || variant >= gemm_num_variants(mode)) {
        set_error("gemm_features: no tile variant %d in family %d", variant, mode);
        return CHAI_ERR_ARG;
    }
    gemm_features(*d, mode, variant, out);
    return CHAI_OK;
}

int chai_hip_gemm_predict_variant(const chai_gemm_desc* d, void* mlp, int* variant, float* predicted, void* stream)
{
    Mlp* N = (Mlp*)mlp;
    if (!d || !N || !variant || N->L.F != CHAI_GEMM_FEATURES) { set_error("gemm_predict_variant: bad arguments"); return CHAI_ERR_ARG; }
    if (!d->M || !d->N || !d->K) { set_error("gemm_predict_variant: empty shape"); return CHAI_ERR_ARG; }
    const int mode = model_mode(*d);
    if (family_slot(mode) < 0) { set_error("gemm_predict_variant: family %d has no tile variants", mode); return CHAI_ERR_ARG; }
    const int nv = gemm_num_variants(mode);
    std::vector<float> X((size_t)nv * CHAI_GEMM_FEATURES), y(nv);
    for (int v = 0; v < nv; v++) gemm_features(*d, mode, v, &X[(size_t)v * CHAI_GEMM_FEATURES]);
    hipStream_t st = as_stream(stream);
    float* dX = nullptr;
    CHAI_HIP_TRY(hipMallocAsync((void**)&dX, sizeof(float) * (X.size() + nv), st));
    float* dy = dX + X.size();
    int rc = CHAI_OK;
    hipError_t e = hipMemcpyAsync(dX, X.data(), sizeof(float) * X.size(), hipMemcpyHostToDevice, st);
    if (e != hipSuccess) rc = hip_fail(e, "gemm_predict_variant: upload");
    if (rc == CHAI_OK) rc = mlp_run(N, dX, nullptr, nv, dy, false, st, nullptr);
    if (rc == CHAI_OK) {
        e = hipMemcpyAsync(y.data(), dy, sizeof(float) * nv, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = hip_fail(e, "gemm_predict_variant: readback");
    }
    (void)hipFreeAsync(dX, st);
    if (rc != CHAI_OK) return rc;
    int best = 0;
    for (int v = 1; v < nv; v++)
        if (y[v] < y[best]) best = v;
    *variant = best;
    if (predicted) std::memcpy(predicted, y.data(), sizeof(float) * nv);
    return chai_hip_gemm_set_variant(d, best);
}

}  // extern "C"
