// ragged.hip -- ragged batches of f32 GEMMs (+ fused epilogue) as one
// persistent split-BF16x3 launch, and the thread-safe batcher around it.
//
// The reference batches only traces with an identical TEA hash, i.e. identical
// shapes (runtime/Scheduler.cpp:48-307, common/VectorTrace.cpp:9-100,
// enqueue/EnqueueMatmulAuto.cpp:57-68); a ragged batch of uneven sizes runs
// there as one kernel (plus one epilogue kernel) per instance.  Here a batch
// of any shapes runs as two launches:
//   1. ragged_split_kernel: every instance's A and B^T become zero-padded
//      bf16 hi/lo panels in k blocks of 32 (one flat grid of 64 x 64 blocks; a
//      block finds its instance by binary search over per-instance block
//      offsets).  Block 0 also clears the tile-queue counters, so the pair can
//      be captured in a hipGraph and replayed without memset nodes.
//   2. ragged_x3_kernel: a persistent grid (CUs x resident blocks) pulls
//      output tiles from eight queues, one per XCD.  Queue q holds the tiles of
//      the instances assigned to it (largest first, greedy least-loaded by
//      tile work), so the blocks of one XCD share A / B panels in that XCD's
//      L2; a block whose queue is empty steals from the next ones.  Each tile
//      is the split-BF16x3 MFMA tile of the dense path (gemm_x3.hpp) with the
//      epilogue fused into its store.
#include "common.hpp"
#include "gemm_device.hpp"
#include "gemm_internal.hpp"
#include "gemm_split.hpp"
#include "gemm_x3.hpp"
#include "gemm_x3_f32.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

namespace chai_hip {

constexpr int kQueues = 8;   // XCDs of MI355X

// device-side instance descriptor
struct RaggedInst {
    const float* A;
    const float* B;
    float* C;
    uint16_t *Ah, *Al, *Bh, *Bl;   // Mp x Kp and Np x Kp panels (k blocks of 32) in the workspace
    uint32_t M, N, K, Kp;
    uint32_t lda, ldb, ldc;
    uint32_t a_blocks;             // split blocks of A: (Mp / 64) * (Kp / 64)
    uint32_t kblocks;              // Kp / 64
    uint32_t Mp, Np;               // panel rows (tile multiples)
    uint32_t pad;
};

struct RaggedEpi {
    int on;
    float scale, shift, lo;
};

// ---------------------------------------------------------------- pre-pass
__global__ __launch_bounds__(256) void ragged_split_kernel(const RaggedInst* __restrict__ inst,
                                                           const uint32_t* __restrict__ first,
                                                           uint32_t count, uint32_t* counters)
{
    __shared__ float tile[64][65];
    const uint32_t b = blockIdx.x;
    if (b == 0 && threadIdx.x < kQueues) counters[threadIdx.x] = 0u;
    // instance i owns split blocks [first[i], first[i+1])
    uint32_t lo = 0, hi = count;   // first[lo] <= b < first[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (first[mid] <= b) lo = mid; else hi = mid;
    }
    const RaggedInst& I = inst[lo];
    uint32_t j = b - first[lo];
    if (j < I.a_blocks) {
        const uint32_t rb = j / I.kblocks, kb = j - rb * I.kblocks;
        split_tile<float, false>(I.A, I.lda, I.M, I.K, I.Kp, I.Mp, I.Ah, I.Al, (uint64_t)rb * 64,
                                 (uint64_t)kb * 64, tile);
    } else {
        j -= I.a_blocks;
        const uint32_t rb = j / I.kblocks, kb = j - rb * I.kblocks;
        // B^T panel element (n, k) = B[k * ldb + n]: transposed read
        split_tile<float, true>(I.B, I.ldb, I.N, I.K, I.Kp, I.Np, I.Bh, I.Bl, (uint64_t)rb * 64,
                                (uint64_t)kb * 64, tile);
    }
}

// ------------------------------------------------------- persistent kernel
// tiles[qmeta[8 + q] + t] = {instance, tm << 16 | tn} for t < qmeta[q]
template <int BM, int BN, int BK, int WGM, int WGN, int MINB>
__global__ __launch_bounds__(WGM * WGN * 64, MINB) void ragged_x3_kernel(
    const RaggedInst* __restrict__ inst, const uint2* __restrict__ tiles,
    const uint32_t* __restrict__ qmeta, uint32_t* counters, RaggedEpi epi)
{
    extern __shared__ u32x4 lds[];
    __shared__ uint32_t s_next;
    const uint32_t home = blockIdx.x % kQueues;
    for (uint32_t qi = 0; qi < kQueues; qi++) {
        const uint32_t q = __builtin_amdgcn_readfirstlane((home + qi) % kQueues);
        const uint32_t qlen = qmeta[q], qoff = qmeta[kQueues + q];
        for (;;) {
            if (threadIdx.x == 0) s_next = atomicAdd(&counters[q], 1u);
            __syncthreads();
            const uint32_t t = __builtin_amdgcn_readfirstlane(s_next);
            __syncthreads();   // every thread has read s_next before thread 0 may rewrite it
            if (t >= qlen) break;
            const uint2 w = tiles[qoff + t];
            const RaggedInst& I = inst[w.x];
            GemmArgs g{};
            g.Ah = I.Ah; g.Al = I.Al; g.Bh = I.Bh; g.Bl = I.Bl;
            g.a_kblk = (uint64_t)I.Mp * 32; g.b_kblk = (uint64_t)I.Np * 32;
            g.Kp = I.Kp;
            g.C = I.C; g.ldc = I.ldc;
            g.M = I.M; g.N = I.N;
            g.epi = epi.on; g.epi_scale = epi.scale; g.epi_shift = epi.shift; g.epi_lo = epi.lo;
            x3_tile_body<BM, BN, BK, WGM, WGN, 0>(g, w.y >> 16, w.y & 0xffffu, 0, lds);   // f32 operands: no zero-panel flags
        }
    }
}

// The f32-fed variants: tiles read A and B directly (gemm_x3_f32.hpp), so the
// plan needs no split pre-pass and no panels; a one-block kernel clears the
// queue counters instead.
template <int BM, int BN, int WGM, int WGN, int MINB>
__global__ __launch_bounds__(WGM * WGN * 64, MINB) void ragged_x3f_kernel(
    const RaggedInst* __restrict__ inst, const uint2* __restrict__ tiles,
    const uint32_t* __restrict__ qmeta, uint32_t* counters, RaggedEpi epi)
{
    extern __shared__ u32x4 lds[];
    __shared__ uint32_t s_next;
    const uint32_t home = blockIdx.x % kQueues;
    for (uint32_t qi = 0; qi < kQueues; qi++) {
        const uint32_t q = __builtin_amdgcn_readfirstlane((home + qi) % kQueues);
        const uint32_t qlen = qmeta[q], qoff = qmeta[kQueues + q];
        for (;;) {
            if (threadIdx.x == 0) s_next = atomicAdd(&counters[q], 1u);
            __syncthreads();
            const uint32_t t = __builtin_amdgcn_readfirstlane(s_next);
            __syncthreads();   // every thread has read s_next before thread 0 may rewrite it
            if (t >= qlen) break;
            const uint2 w = tiles[qoff + t];
            const RaggedInst& I = inst[w.x];
            GemmArgs g{};
            g.Af = I.A; g.Bf = I.B; g.lda = I.lda; g.ldb = I.ldb;
            g.K = I.K; g.Kp = I.Kp;
            g.C = I.C; g.ldc = I.ldc;
            g.M = I.M; g.N = I.N;
            g.epi = epi.on; g.epi_scale = epi.scale; g.epi_shift = epi.shift; g.epi_lo = epi.lo;
            x3f_tile<BM, BN, WGM, WGN, MINB == 1>(g, w.y >> 16, w.y & 0xffffu, lds);
        }
    }
}

__global__ void ragged_clear_counters(uint32_t* counters)
{
    if (threadIdx.x < kQueues) counters[threadIdx.x] = 0u;
}

struct RaggedVariant {
    int BM, BN, BK;
    size_t lds;
    int threads;
    void (*kernel)(const RaggedInst*, const uint2*, const uint32_t*, uint32_t*, RaggedEpi);
    bool fused;    // f32-fed tiles: no split pre-pass
};

#define CHAI_RAGGEDV(BM, BN, BK, WGM, WGN, MINB)                                              \
    {BM, BN, BK, X3Shape<BM, BN, BK, WGM, WGN>::LDS_BYTES, WGM * WGN * 64,                    \
     ragged_x3_kernel<BM, BN, BK, WGM, WGN, MINB>, false},
#define CHAI_RAGGEDF(BM, BN, WGM, WGN, MINB)                                                  \
    {BM, BN, 32, X3Shape<BM, BN, 32, WGM, WGN>::LDS_BYTES, WGM * WGN * 64,                    \
     ragged_x3f_kernel<BM, BN, WGM, WGN, MINB>, true},

// variants 0..7: the panel tiles of the dense path (split pre-pass first);
// 8..9: the f32-fed tiles
static const RaggedVariant kRaggedVariants[] = {CHAI_X3_VARIANTS(CHAI_RAGGEDV) CHAI_X3F_VARIANTS(CHAI_RAGGEDF)};
static constexpr int kNumRagged = sizeof(kRaggedVariants) / sizeof(kRaggedVariants[0]);
static int g_ragged_tuned = -1;
static std::mutex g_ragged_mu;

static uint64_t round_up(uint64_t x, uint64_t m) { return (x + m - 1) / m * m; }

struct RaggedPlan {
    int variant = 0;
    uint32_t count = 0;          // instances with a non-empty output
    uint32_t split_blocks = 0;
    uint32_t ntiles = 0;
    uint32_t grid = 0;
    RaggedEpi epi{};
    void* mem = nullptr;         // one allocation: descriptors, offsets, tiles, counters, panels
    size_t mem_bytes = 0;
    RaggedInst* d_inst = nullptr;
    uint32_t* d_first = nullptr;
    uint2* d_tiles = nullptr;
    uint32_t* d_qmeta = nullptr;
    uint32_t* d_counters = nullptr;
    uint64_t useful_flops = 0, padded_flops = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t graph_stream = nullptr;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    bool timed = false;

    ~RaggedPlan()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (mem) (void)hipFree(mem);
    }
};

static int ragged_build(const chai_ragged_item* items, size_t count, const chai_epilogue* epi,
                        int variant, RaggedPlan** out)
{
    *out = nullptr;
    if (count && !items) { set_error("ragged_plan: null items"); return CHAI_ERR_ARG; }
    if (epi && epi->kind != CHAI_EPI_NONE && epi->kind != CHAI_EPI_AFFINE_MAX) {
        set_error("ragged_plan: unknown epilogue kind %d", epi->kind);
        return CHAI_ERR_ARG;
    }
    if (variant < 0) {
        std::lock_guard<std::mutex> lk(g_ragged_mu);
        variant = g_ragged_tuned >= 0 ? g_ragged_tuned : 0;
    }
    if (variant >= kNumRagged) { set_error("ragged_plan: variant %d out of range", variant); return CHAI_ERR_ARG; }
    const RaggedVariant& v = kRaggedVariants[variant];
    auto* P = new RaggedPlan;
    P->variant = variant;
    if (epi && epi->kind == CHAI_EPI_AFFINE_MAX) P->epi = RaggedEpi{1, epi->scale, epi->shift, epi->lo};

    // validate, keep instances with work
    std::vector<size_t> idx;
    for (size_t i = 0; i < count; i++) {
        const chai_ragged_item& it = items[i];
        if (it.M == 0 || it.N == 0) continue;
        if (!it.A || !it.B || !it.C || it.K == 0) {
            set_error("ragged_plan: item %zu has a null operand or K == 0", i);
            delete P;
            return CHAI_ERR_ARG;
        }
        if (it.lda < it.K || it.ldb < it.N || it.ldc < it.N) {
            set_error("ragged_plan: item %zu has a leading dimension below its row length", i);
            delete P;
            return CHAI_ERR_ARG;
        }
        const uint64_t lim = 1ull << 30;
        if (it.M >= lim || it.N >= lim || it.K >= lim || it.lda >= lim || it.ldb >= lim || it.ldc >= lim ||
            round_up(it.M, v.BM) / v.BM > 0xffff || round_up(it.N, v.BN) / v.BN > 0xffff) {
            set_error("ragged_plan: item %zu is too large for a ragged batch", i);
            delete P;
            return CHAI_ERR_ARG;
        }
        // the panel tiles address a panel's k block pair with 32-bit offsets
        if (!v.fused && (round_up(it.M, v.BM) > kX3MaxRows || round_up(it.N, v.BN) > kX3MaxRows)) {
            set_error("ragged_plan: item %zu: the panel tile variants take at most 2^25 padded rows / columns", i);
            delete P;
            return CHAI_ERR_ARG;
        }
        // the f32-fed tiles address a wave's rows through 31-bit buffer offsets
        if (v.fused && (it.lda >= (1ull << 23) || it.ldb >= (1ull << 23))) {
            set_error("ragged_plan: item %zu: the f32-fed tile variants need row strides below 2^23", i);
            delete P;
            return CHAI_ERR_ARG;
        }
        idx.push_back(i);
    }
    const size_t n = idx.size();
    P->count = (uint32_t)n;
    if (n == 0) { *out = P; return CHAI_OK; }

    // largest instance first (work = tiles x Kp)
    struct Shape { uint64_t Mp, Np, Kp, tm, tn, work; };
    std::vector<Shape> sh(n);
    for (size_t j = 0; j < n; j++) {
        const chai_ragged_item& it = items[idx[j]];
        Shape& s = sh[j];
        s.Mp = round_up(it.M, v.BM); s.Np = round_up(it.N, v.BN); s.Kp = round_up(it.K, 64);
        s.tm = s.Mp / v.BM; s.tn = s.Np / v.BN;
        s.work = s.tm * s.tn * s.Kp;
    }
    std::vector<uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return sh[a].work > sh[b].work; });

    // layout of the single allocation
    size_t off = 0;
    auto take = [&](size_t bytes, size_t align) { off = round_up(off, align); const size_t o = off; off += bytes; return o; };
    const size_t o_inst = take(n * sizeof(RaggedInst), 256);
    const size_t o_first = take((n + 1) * sizeof(uint32_t), 256);
    uint64_t ntiles = 0, nsplit = 0;
    for (size_t j = 0; j < n; j++) {
        ntiles += sh[j].tm * sh[j].tn;
        nsplit += (sh[j].Mp / 64 + sh[j].Np / 64) * (sh[j].Kp / 64);
    }
    if (ntiles > 0xffffffffull || nsplit > 0x7fffffffull) {
        set_error("ragged_plan: batch too large (%llu tiles)", (unsigned long long)ntiles);
        delete P;
        return CHAI_ERR_ARG;
    }
    const size_t o_tiles = take(ntiles * sizeof(uint2), 256);
    const size_t o_qmeta = take(2 * kQueues * sizeof(uint32_t), 256);
    const size_t o_cnt = take(kQueues * sizeof(uint32_t), 256);
    std::vector<size_t> o_panel(n, 0);
    if (!v.fused)
        for (size_t j = 0; j < n; j++) o_panel[j] = take(2 * sizeof(uint16_t) * (sh[j].Mp + sh[j].Np) * sh[j].Kp, 256);
    P->mem_bytes = off;

    // host images
    std::vector<RaggedInst> hinst(n);
    std::vector<uint32_t> hfirst(n + 1);
    uint32_t acc = 0;
    for (size_t j = 0; j < n; j++) {
        const chai_ragged_item& it = items[idx[j]];
        const Shape& s = sh[j];
        RaggedInst& I = hinst[j];
        std::memset(&I, 0, sizeof I);
        I.A = it.A; I.B = it.B; I.C = it.C;
        I.M = (uint32_t)it.M; I.N = (uint32_t)it.N; I.K = (uint32_t)it.K; I.Kp = (uint32_t)s.Kp;
        I.lda = (uint32_t)it.lda; I.ldb = (uint32_t)it.ldb; I.ldc = (uint32_t)it.ldc;
        I.kblocks = (uint32_t)(s.Kp / 64);
        I.Mp = (uint32_t)s.Mp; I.Np = (uint32_t)s.Np;
        I.a_blocks = (uint32_t)(s.Mp / 64) * I.kblocks;
        hfirst[j] = acc;
        acc += I.a_blocks + (uint32_t)(s.Np / 64) * I.kblocks;
        P->useful_flops += 2ull * it.M * it.N * it.K;
        P->padded_flops += 2ull * s.tm * s.tn * v.BM * v.BN * s.Kp;
    }
    hfirst[n] = acc;
    P->split_blocks = v.fused ? 0 : acc;

    // greedy LPT assignment of instances to the per-XCD queues
    std::vector<std::vector<uint2>> q(kQueues);
    uint64_t load[kQueues] = {};
    for (uint32_t j : order) {
        int best = 0;
        for (int k = 1; k < kQueues; k++)
            if (load[k] < load[best]) best = k;
        load[best] += sh[j].work;
        for (uint64_t tm = 0; tm < sh[j].tm; tm++)
            for (uint64_t tn = 0; tn < sh[j].tn; tn++) q[best].push_back(make_uint2(j, (uint32_t)(tm << 16 | tn)));
    }
    std::vector<uint2> htiles;
    htiles.reserve(ntiles);
    uint32_t hq[2 * kQueues];
    for (int k = 0; k < kQueues; k++) {
        hq[k] = (uint32_t)q[k].size();
        hq[kQueues + k] = (uint32_t)htiles.size();
        htiles.insert(htiles.end(), q[k].begin(), q[k].end());
    }
    P->ntiles = (uint32_t)ntiles;

    hipError_t e = hipMalloc(&P->mem, P->mem_bytes);
    if (e != hipSuccess) { P->mem = nullptr; delete P; return hip_fail(e, "ragged_plan: workspace"); }
    char* base = (char*)P->mem;
    for (size_t j = 0; j < n; j++) {
        if (v.fused) continue;   // no panels
        uint16_t* p = (uint16_t*)(base + o_panel[j]);
        const uint64_t a = sh[j].Mp * sh[j].Kp, b = sh[j].Np * sh[j].Kp;
        hinst[j].Ah = p; hinst[j].Al = p + a; hinst[j].Bh = p + 2 * a; hinst[j].Bl = p + 2 * a + b;
    }
    P->d_inst = (RaggedInst*)(base + o_inst);
    P->d_first = (uint32_t*)(base + o_first);
    P->d_tiles = (uint2*)(base + o_tiles);
    P->d_qmeta = (uint32_t*)(base + o_qmeta);
    P->d_counters = (uint32_t*)(base + o_cnt);
    e = hipMemcpy(P->d_inst, hinst.data(), n * sizeof(RaggedInst), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(P->d_first, hfirst.data(), (n + 1) * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(P->d_tiles, htiles.data(), ntiles * sizeof(uint2), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(P->d_qmeta, hq, sizeof hq, hipMemcpyHostToDevice);
    if (e != hipSuccess) { delete P; return hip_fail(e, "ragged_plan: upload"); }

    // persistent grid: every resident block slot, a multiple of the queue count
    static bool attr_set[kNumRagged] = {};
    if (!attr_set[variant]) {
        e = hipFuncSetAttribute((const void*)v.kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)v.lds);
        if (e != hipSuccess) { delete P; return hip_fail(e, "ragged_plan: LDS attribute"); }
        attr_set[variant] = true;
    }
    int dev = 0, cus = 0, per_cu = 0;
    e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, v.kernel, v.threads, v.lds);
    if (e != hipSuccess) { delete P; return hip_fail(e, "ragged_plan: occupancy"); }
    uint64_t grid = (uint64_t)std::max(cus, 1) * (uint64_t)std::max(per_cu, 1);
    grid = std::max<uint64_t>(kQueues, grid / kQueues * kQueues);
    P->grid = (uint32_t)std::min<uint64_t>(grid, round_up(ntiles, kQueues));
    *out = P;
    return CHAI_OK;
}

static int ragged_launch(RaggedPlan* P, hipStream_t st, bool timed)
{
    if (P->count == 0) return CHAI_OK;
    const RaggedVariant& v = kRaggedVariants[P->variant];
    if (timed) CHAI_HIP_TRY(hipEventRecord(P->ev[0], st));
    if (v.fused) {
        ragged_clear_counters<<<1, 64, 0, st>>>(P->d_counters);
        CHAI_LAUNCH_CHECK("ragged_clear_counters");
    } else {
        ragged_split_kernel<<<P->split_blocks, 256, 0, st>>>(P->d_inst, P->d_first, P->count, P->d_counters);
        CHAI_LAUNCH_CHECK("ragged_split_kernel");
    }
    if (timed) CHAI_HIP_TRY(hipEventRecord(P->ev[1], st));
    hipLaunchKernelGGL(v.kernel, dim3(P->grid), dim3(v.threads), v.lds, st, (const RaggedInst*)P->d_inst,
                       (const uint2*)P->d_tiles, (const uint32_t*)P->d_qmeta, P->d_counters, P->epi);
    CHAI_LAUNCH_CHECK("ragged_x3_kernel");
    if (timed) CHAI_HIP_TRY(hipEventRecord(P->ev[2], st));
    return CHAI_OK;
}

static int ragged_run(RaggedPlan* P, int use_graph, hipStream_t st)
{
    if (P->count == 0) return CHAI_OK;
    if (!use_graph) {
        for (auto& e : P->ev)
            if (!e) CHAI_HIP_TRY(hipEventCreate(&e));
        const int rc = ragged_launch(P, st, true);
        P->timed = rc == CHAI_OK;
        return rc;
    }
    P->timed = false;
    if (!P->exec) {
        // capture on a private stream (the legacy null stream cannot be captured)
        hipStream_t cs = nullptr;
        CHAI_HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        int rc = e == hipSuccess ? ragged_launch(P, cs, false) : hip_fail(e, "ragged: begin capture");
        hipGraph_t g = nullptr;
        e = hipStreamEndCapture(cs, &g);
        if (rc == CHAI_OK && e != hipSuccess) rc = hip_fail(e, "ragged: end capture");
        if (rc == CHAI_OK) {
            e = hipGraphInstantiate(&P->exec, g, nullptr, nullptr, 0);
            if (e != hipSuccess) { P->exec = nullptr; rc = hip_fail(e, "ragged: graph instantiate"); }
        }
        if (rc == CHAI_OK) P->graph = g;
        else if (g) (void)hipGraphDestroy(g);
        (void)hipStreamDestroy(cs);
        if (rc != CHAI_OK) return rc;
    }
    CHAI_HIP_TRY(hipGraphLaunch(P->exec, st));
    return CHAI_OK;
}

}  // namespace chai_hip

using namespace chai_hip;

extern "C" {

int chai_hip_ragged_plan(const chai_ragged_item* items, size_t count, const chai_epilogue* epi,
                         int variant, void** plan)
{
    if (!plan) { set_error("ragged_plan: null out"); return CHAI_ERR_ARG; }
    RaggedPlan* P = nullptr;
    const int rc = ragged_build(items, count, epi, variant, &P);
    *plan = P;
    return rc;
}

int chai_hip_ragged_run(void* plan, int use_graph, void* stream)
{
    if (!plan) { set_error("ragged_run: null plan"); return CHAI_ERR_ARG; }
    return ragged_run((RaggedPlan*)plan, use_graph, as_stream(stream));
}

int chai_hip_ragged_info(void* plan, uint64_t* info)
{
    if (!plan || !info) { set_error("ragged_info: null argument"); return CHAI_ERR_ARG; }
    const RaggedPlan* P = (const RaggedPlan*)plan;
    info[0] = P->ntiles; info[1] = P->count ? P->grid : 0; info[2] = P->mem_bytes;
    info[3] = P->useful_flops; info[4] = P->padded_flops; info[5] = (uint64_t)P->variant;
    info[6] = P->split_blocks;
    return CHAI_OK;
}

int chai_hip_ragged_last_timing(void* plan, float* split_ms, float* core_ms)
{
    RaggedPlan* P = (RaggedPlan*)plan;
    if (!P || !P->timed) { set_error("ragged_last_timing: no timed (non-graph) run on this plan"); return CHAI_ERR_STATE; }
    CHAI_HIP_TRY(hipEventSynchronize(P->ev[2]));
    float a = 0.f, b = 0.f;
    CHAI_HIP_TRY(hipEventElapsedTime(&a, P->ev[0], P->ev[1]));
    CHAI_HIP_TRY(hipEventElapsedTime(&b, P->ev[1], P->ev[2]));
    if (split_ms) *split_ms = a;
    if (core_ms) *core_ms = b;
    return CHAI_OK;
}

int chai_hip_ragged_release(void* plan)
{
    RaggedPlan* P = (RaggedPlan*)plan;
    if (P && (P->exec || P->timed)) (void)hipDeviceSynchronize();   // nothing of the plan may be in flight
    delete P;
    return CHAI_OK;
}

int chai_hip_ragged_num_variants(int* n)
{
    if (!n) { set_error("ragged_num_variants: null output"); return CHAI_ERR_ARG; }
    *n = kNumRagged;
    return CHAI_OK;
}

int chai_hip_ragged_tune(const chai_ragged_item* items, size_t count, const chai_epilogue* epi,
                         int trials, int* best_variant, float* best_ms, void* stream)
{
    if (trials < 1) { set_error("ragged_tune: trials < 1"); return CHAI_ERR_ARG; }
    hipStream_t st = as_stream(stream);
    int bv = -1;
    float bt = 1e30f;
    for (int v = 0; v < kNumRagged; v++) {
        RaggedPlan* P = nullptr;
        int rc = ragged_build(items, count, epi, v, &P);
        if (rc == CHAI_ERR_ARG && v > 0 && kRaggedVariants[v].fused) continue;   // strides too large for it
        if (rc != CHAI_OK) return rc;
        float best = 1e30f;
        for (int t = 0; t <= trials && rc == CHAI_OK; t++) {   // run 0 is the warm-up
            rc = ragged_run(P, 0, st);
            float a = 0.f, b = 0.f;
            if (rc == CHAI_OK) rc = chai_hip_ragged_last_timing(P, &a, &b);
            if (rc == CHAI_OK && t > 0 && a + b < best) best = a + b;
        }
        chai_hip_ragged_release(P);
        if (rc != CHAI_OK) return rc;
        if (best < bt) { bt = best; bv = v; }
    }
    {
        std::lock_guard<std::mutex> lk(g_ragged_mu);
        g_ragged_tuned = bv;
    }
    if (best_variant) *best_variant = bv;
    if (best_ms) *best_ms = bt;
    return CHAI_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ batcher
namespace {

struct Batcher {
    std::mutex mu;                 // pending items, results
    std::condition_variable cv;
    std::mutex run_mu;             // one batch on the device at a time
    size_t expected = 0;
    int timeout_ms = 0;            // 0: wait for a full batch
    chai_epilogue epi{};
    int variant = -1, use_graph = 1;
    hipStream_t stream = nullptr;
    std::vector<chai_ragged_item> pending;
    std::vector<hipEvent_t> pending_ev;   // per pending item: its producer's event
    uint64_t gen = 0;              // generation of the pending batch
    std::vector<int> result;       // status of each finished generation
    // cached plan of the last batch, keyed by its items
    std::vector<chai_ragged_item> plan_items;
    RaggedPlan* plan = nullptr;
    uint64_t batches = 0, items = 0, plans = 0, replays = 0;

    ~Batcher()
    {
        if (plan) chai_hip_ragged_release(plan);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

bool same_items(const std::vector<chai_ragged_item>& a, const std::vector<chai_ragged_item>& b)
{
    return a.size() == b.size() &&
           (a.empty() || std::memcmp(a.data(), b.data(), a.size() * sizeof(chai_ragged_item)) == 0);
}

// Runs one batch; called with run_mu held and mu released.  The batch's
// stream first waits for the event each submitter recorded on the stream
// that produced its operands.
int run_batch(Batcher* B, std::vector<chai_ragged_item>& items, const std::vector<hipEvent_t>& events)
{
    int rc = CHAI_OK;
    hipEvent_t last = nullptr;
    for (hipEvent_t ev : events)
        if (ev != last) { CHAI_HIP_TRY(hipStreamWaitEvent(B->stream, ev, 0)); last = ev; }
    if (!B->plan || !same_items(items, B->plan_items)) {
        if (B->plan) { chai_hip_ragged_release(B->plan); B->plan = nullptr; }
        if (B->variant < 0) {
            int bv = 0;
            rc = chai_hip_ragged_tune(items.data(), items.size(), &B->epi, 2, &bv, nullptr, B->stream);
            if (rc != CHAI_OK) return rc;
            B->variant = bv;
        }
        RaggedPlan* P = nullptr;
        rc = ragged_build(items.data(), items.size(), &B->epi, B->variant, &P);
        if (rc != CHAI_OK) return rc;
        B->plan = P;
        B->plan_items = items;
        B->plans++;
    } else if (B->use_graph) {
        B->replays++;
    }
    rc = ragged_run(B->plan, B->use_graph, B->stream);
    if (rc == CHAI_OK) CHAI_HIP_TRY(hipStreamSynchronize(B->stream));
    B->batches++;
    B->items += items.size();
    return rc;
}

// Takes the pending batch (mu held on entry and exit), runs it, publishes the status.
int flush_locked(Batcher* B, std::unique_lock<std::mutex>& lk)
{
    std::vector<chai_ragged_item> items;
    std::vector<hipEvent_t> events;
    items.swap(B->pending);
    events.swap(B->pending_ev);
    const uint64_t g = B->gen++;
    lk.unlock();
    int rc;
    {
        std::lock_guard<std::mutex> rl(B->run_mu);
        rc = run_batch(B, items, events);
    }
    lk.lock();
    if (B->result.size() <= g) B->result.resize(g + 1, 1);
    B->result[g] = rc;
    B->cv.notify_all();
    return rc;
}

}  // namespace

extern "C" {

int chai_hip_batcher_create(size_t expected, const chai_epilogue* epi, int variant, int use_graph,
                            int timeout_ms, void** batcher)
{
    if (!batcher || expected == 0 || timeout_ms < 0) { set_error("batcher_create: bad arguments"); return CHAI_ERR_ARG; }
    if (variant >= kNumRagged) { set_error("batcher_create: variant %d out of range", variant); return CHAI_ERR_ARG; }
    auto* B = new Batcher;
    B->expected = expected;
    if (epi) B->epi = *epi;
    B->variant = variant;
    B->use_graph = use_graph;
    B->timeout_ms = timeout_ms;
    hipError_t e = hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { B->stream = nullptr; delete B; return hip_fail(e, "batcher_create: stream"); }
    *batcher = B;
    return CHAI_OK;
}

int chai_hip_batcher_submit(void* batcher, const chai_ragged_item* items, size_t count, void* stream)
{
    if (!batcher || (count && !items)) { set_error("batcher_submit: null argument"); return CHAI_ERR_ARG; }
    auto* B = (Batcher*)batcher;
    if (count == 0) return CHAI_OK;
    // order the batch after the work already queued on the producer's stream
    hipEvent_t ev = nullptr;
    CHAI_HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    hipError_t e = hipEventRecord(ev, as_stream(stream));
    if (e != hipSuccess) { (void)hipEventDestroy(ev); return hip_fail(e, "batcher_submit: event record"); }
    std::unique_lock<std::mutex> lk(B->mu);
    const uint64_t first = B->gen;
    uint64_t last = first;
    for (size_t i = 0; i < count; i++) {
        B->pending.push_back(items[i]);
        B->pending_ev.push_back(ev);
        last = B->gen;
        if (B->pending.size() >= B->expected) flush_locked(B, lk);
    }
    auto done = [&] {
        for (uint64_t g = first; g <= last; g++)
            if (B->result.size() <= g || B->result[g] == 1) return false;
        return true;
    };
    while (!done()) {
        if (B->timeout_ms == 0) {
            B->cv.wait(lk);
        } else if (!B->cv.wait_for(lk, std::chrono::milliseconds(B->timeout_ms), done)) {
            // nobody filled the batch in time: run what has been gathered
            if (B->gen == last && !B->pending.empty()) flush_locked(B, lk);
        }
    }
    int rc = CHAI_OK;
    for (uint64_t g = first; g <= last; g++)
        if (B->result[g] != CHAI_OK) rc = B->result[g];
    lk.unlock();
    // every batch holding these items has finished (run_batch synchronises)
    (void)hipEventDestroy(ev);
    return rc;
}

int chai_hip_batcher_flush(void* batcher)
{
    if (!batcher) { set_error("batcher_flush: null batcher"); return CHAI_ERR_ARG; }
    auto* B = (Batcher*)batcher;
    std::unique_lock<std::mutex> lk(B->mu);
    if (B->pending.empty()) return CHAI_OK;
    return flush_locked(B, lk);
}

int chai_hip_batcher_stats(void* batcher, uint64_t* stats)
{
    if (!batcher || !stats) { set_error("batcher_stats: null argument"); return CHAI_ERR_ARG; }
    auto* B = (Batcher*)batcher;
    std::lock_guard<std::mutex> rl(B->run_mu);
    stats[0] = B->batches; stats[1] = B->items; stats[2] = B->plans; stats[3] = B->replays;
    stats[4] = (uint64_t)(int64_t)B->variant;
    return CHAI_OK;
}

int chai_hip_batcher_destroy(void* batcher)
{
    auto* B = (Batcher*)batcher;
    if (!B) return CHAI_OK;
    {
        std::lock_guard<std::mutex> lk(B->mu);
        if (!B->pending.empty()) {
            set_error("batcher_destroy: %zu items still pending (flush first)", B->pending.size());
            return CHAI_ERR_STATE;
        }
    }
    { std::lock_guard<std::mutex> rl(B->run_mu); }
    delete B;
    return CHAI_OK;
}

}  // extern "C"
