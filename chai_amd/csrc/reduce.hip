// reduce.hip -- sum / mean / dot over each slot, FP64 cross-block combine.
//
// The reference reduces with one work item looping over the whole array in
// element precision (enqueue/EnqueueReduce.cpp:45-103); its interpreter
// accumulates in double (interp/InterpAccum.cpp:34-63).  Here every thread
// streams 16-byte vectors from HBM and accumulates in FP64 (integers in
// int64, exact), waves reduce with cross-lane shuffles, blocks write one FP64
// partial each, and a second single-block pass combines the partials of each
// slot in a fixed order -- so results are run-to-run deterministic and agree
// with the interpreter's double accumulation to within its own rounding.
#include "common.hpp"
#include "reduce_device.hpp"

#include <mutex>

namespace chai_hip {

constexpr int kRedThreads = 256;

template <typename T>
__device__ __forceinline__ double as_acc(T v) { return (double)v; }

// Integer inputs accumulate exactly in int64 and are returned as double.
template <typename T, bool IS_INT>
__global__ __launch_bounds__(kRedThreads) void reduce_partial(const T* __restrict__ in, uint64_t n,
                                                              uint32_t blocks_per_slot,
                                                              double* __restrict__ partial)
{
    const uint32_t slot = blockIdx.y;
    const T* p = in + (uint64_t)slot * n;
    const uint64_t tid = (uint64_t)blockIdx.x * kRedThreads + threadIdx.x;
    const uint64_t stride = (uint64_t)blocks_per_slot * kRedThreads;
    double accd = 0.0;
    long long acci = 0;
    constexpr int VEC = 16 / sizeof(T);
    const bool vec_ok = ((uintptr_t)p % 16) == 0;
    uint64_t done = 0;
    if (vec_ok) {
        typedef T TV __attribute__((ext_vector_type(VEC)));
        const TV* pv = reinterpret_cast<const TV*>(p);
        const uint64_t nv = n / VEC;
        for (uint64_t i = tid; i < nv; i += stride) {
            TV v = __builtin_nontemporal_load(pv + i);
#pragma unroll
            for (int j = 0; j < VEC; j++) {
                if (IS_INT) acci += (long long)v[j];
                else accd += (double)v[j];
            }
        }
        done = nv * VEC;
    }
    for (uint64_t i = done + tid; i < n; i += stride) {
        if (IS_INT) acci += (long long)p[i];
        else accd += (double)p[i];
    }
    double r;
    if (IS_INT) r = (double)block_sum_i64<kRedThreads>(acci);
    else r = block_sum_f64<kRedThreads>(accd);
    if (threadIdx.x == 0) partial[(uint64_t)slot * blocks_per_slot + blockIdx.x] = r;
}

// dot(a, b) with the product formed in the promoted type (float*float -> float).
template <typename TA, typename TB, typename TP>
__global__ __launch_bounds__(kRedThreads) void dot_partial(const TA* __restrict__ a,
                                                           const TB* __restrict__ b, uint64_t n,
                                                           uint32_t blocks_per_slot,
                                                           double* __restrict__ partial)
{
    const uint32_t slot = blockIdx.y;
    const TA* pa = a + (uint64_t)slot * n;
    const TB* pb = b + (uint64_t)slot * n;
    const uint64_t tid = (uint64_t)blockIdx.x * kRedThreads + threadIdx.x;
    const uint64_t stride = (uint64_t)blocks_per_slot * kRedThreads;
    double acc = 0.0;
    for (uint64_t i = tid; i < n; i += stride) {
        const TP prod = (TP)pa[i] * (TP)pb[i];
        acc += (double)prod;
    }
    const double r = block_sum_f64<kRedThreads>(acc);
    if (threadIdx.x == 0) partial[(uint64_t)slot * blocks_per_slot + blockIdx.x] = r;
}

// The result in the element type, converted the way the interpreter's x86
// build converts its double accumulator (cvttsd2si): to int32 a value outside
// [-2^31, 2^31) gives INT_MIN; to uint32 the conversion goes through int64 and
// keeps the low 32 bits.
template <typename T>
__device__ __forceinline__ T from_double(double v)
{
    if constexpr (std::is_same<T, uint32_t>::value)
        return (v != v || v >= 0x1p63 || v < -0x1p63) ? 0u : (uint32_t)(unsigned long long)(long long)v;
    else if constexpr (std::is_same<T, int32_t>::value)
        return (v != v || v >= 0x1p31 || v < -0x1p31) ? INT32_MIN : (int32_t)v;
    else return (T)v;
}

// Fixed-order combine of one slot's partials (one block per slot).
template <typename T>
__global__ __launch_bounds__(kRedThreads) void reduce_final(const double* __restrict__ partial,
                                                            uint32_t blocks_per_slot, double divisor,
                                                            T* __restrict__ out,
                                                            double* __restrict__ out_f64)
{
    const uint32_t slot = blockIdx.x;
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < blocks_per_slot; i += kRedThreads)
        acc += partial[(uint64_t)slot * blocks_per_slot + i];
    double r = block_sum_f64<kRedThreads>(acc);
    if (threadIdx.x == 0) {
        if (divisor != 1.0) r = r / divisor;
        out[slot] = from_double<T>(r);
        if (out_f64) out_f64[slot] = r;
    }
}

static uint32_t pick_blocks(uint64_t n, size_t slots)
{
    // enough blocks to cover 256 CUs x 8 waves when there is one slot; each
    // thread should see at least ~64 bytes of input.
    uint64_t want = (n + kRedThreads * 16 - 1) / (kRedThreads * 16);
    uint64_t cap = 2048 / (slots < 8 ? slots : 8);
    if (want > cap) want = cap;
    if (want < 1) want = 1;
    return (uint32_t)want;
}

// The two launches of one reduction over a caller-provided partials buffer
// of at least bps * slots doubles.
template <typename T>
static hipError_t launch_reduce(int op, const void* in, size_t n, size_t slots, uint32_t bps,
                                double* partial, void* out, double* out_f64, hipStream_t st)
{
    dim3 grid(bps, (unsigned)slots);
    reduce_partial<T, std::is_integral<T>::value><<<grid, kRedThreads, 0, st>>>((const T*)in, n, bps, partial);
    const double div = op == CHAI_OP_MEAN ? (double)n : 1.0;
    reduce_final<T><<<(unsigned)slots, kRedThreads, 0, st>>>(partial, bps, div, (T*)out, out_f64);
    return hipGetLastError();
}

template <typename T>
static int run_reduce(int op, const void* in, size_t n, size_t slots, void* out, double* out_f64,
                      hipStream_t st)
{
    const uint32_t bps = pick_blocks(n, slots);
    double* partial = nullptr;
    CHAI_HIP_TRY(hipMallocAsync((void**)&partial, sizeof(double) * bps * slots, st));
    hipError_t e = launch_reduce<T>(op, in, n, slots, bps, partial, out, out_f64, st);
    CHAI_HIP_TRY(hipFreeAsync(partial, st));
    if (e != hipSuccess) return hip_fail(e, "reduce launch");
    return CHAI_OK;
}

// ---------------------------------------------------------------------------
// Reduction without allocation.  The block partials live in a file-scope
// device array (one copy per GPU, held until the process ends).  Host code
// never passes the array's name to a kernel: on the host that name is only a
// stand-in, so each device's address is looked up once with
// hipGetSymbolAddress (at chai_hip_init, or on the first call on a device
// that init did not set up) and kept per device.  Every partial is written by
// reduce_partial before reduce_final reads it, so nothing needs clearing
// between calls; calls that overlap on one GPU share the array and race.
constexpr size_t kNoAllocPartials = 65536;   // 512 KiB of FP64 partials
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kNoAllocPartials];

static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices];

int reduce_scratch_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) {
        set_error("reduce scratch: device %d beyond the %d supported", device, kMaxDevices);
        return CHAI_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return CHAI_OK;
    int cur = -1;
    CHAI_HIP_TRY(hipGetDevice(&cur));
    if (cur != device) CHAI_HIP_TRY(hipSetDevice(device));
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(reduce scratch)");
    g_scratch_ptr[device] = static_cast<double*>(ptr);
    return CHAI_OK;
}

static int scratch_for_current_device(double** out)
{
    int dev = -1;
    CHAI_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("reduce scratch: device %d beyond the %d supported", dev, kMaxDevices);
        return CHAI_ERR_ARG;
    }
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (g_scratch_ptr[dev]) { *out = g_scratch_ptr[dev]; return CHAI_OK; }
    }
    int rc = reduce_scratch_setup(dev);
    if (rc != CHAI_OK) return rc;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    *out = g_scratch_ptr[dev];
    return CHAI_OK;
}

template <typename T>
static int run_reduce_noalloc(int op, const void* in, size_t n, size_t slots, void* out,
                              double* out_f64, hipStream_t st)
{
    const uint32_t bps = pick_blocks(n, slots);
    if ((uint64_t)bps * slots > kNoAllocPartials) {
        set_error("reduce_noalloc: %zu slots of %zu elements need %llu FP64 partials; the scratch "
                  "array holds %zu (up to 256 slots always fit)", slots, n,
                  (unsigned long long)bps * slots, kNoAllocPartials);
        return CHAI_ERR_ARG;
    }
    double* partial = nullptr;
    int rc = scratch_for_current_device(&partial);
    if (rc != CHAI_OK) return rc;
    hipError_t e = launch_reduce<T>(op, in, n, slots, bps, partial, out, out_f64, st);
    if (e != hipSuccess) return hip_fail(e, "reduce_noalloc launch");
    return CHAI_OK;
}

}  // namespace chai_hip

using namespace chai_hip;

static int check_reduce_args(const char* who, int op, int prec, const void* in, size_t n, size_t slots,
                             const void* out)
{
    if (op != CHAI_OP_SUM && op != CHAI_OP_MEAN) { set_error("%s: op %d is not sum/mean", who, op); return CHAI_ERR_ARG; }
    if (!valid_prec(prec)) { set_error("%s: bad precision %d", who, prec); return CHAI_ERR_ARG; }
    if (slots == 0) return CHAI_OK;
    if (n == 0) { set_error("%s: empty array", who); return CHAI_ERR_ARG; }
    if (slots > 65535) { set_error("%s: at most 65535 slots per call", who); return CHAI_ERR_ARG; }
    if (!in || !out) { set_error("%s: null pointer", who); return CHAI_ERR_ARG; }
    return CHAI_OK;
}

extern "C" int chai_hip_reduce(int op, int prec, const void* in, size_t n, size_t slots, void* out,
                               double* out_f64, void* stream)
{
    int rc = check_reduce_args("reduce", op, prec, in, n, slots, out);
    if (rc != CHAI_OK || slots == 0) return rc;
    hipStream_t st = as_stream(stream);
    switch (prec) {
        case CHAI_U32: return run_reduce<uint32_t>(op, in, n, slots, out, out_f64, st);
        case CHAI_I32: return run_reduce<int32_t>(op, in, n, slots, out, out_f64, st);
        case CHAI_F32: return run_reduce<float>(op, in, n, slots, out, out_f64, st);
        default: return run_reduce<double>(op, in, n, slots, out, out_f64, st);
    }
}

extern "C" int chai_hip_reduce_noalloc(int op, int prec, const void* in, size_t n, size_t slots, void* out,
                                       double* out_f64, void* stream)
{
    int rc = check_reduce_args("reduce_noalloc", op, prec, in, n, slots, out);
    if (rc != CHAI_OK || slots == 0) return rc;
    hipStream_t st = as_stream(stream);
    switch (prec) {
        case CHAI_U32: return run_reduce_noalloc<uint32_t>(op, in, n, slots, out, out_f64, st);
        case CHAI_I32: return run_reduce_noalloc<int32_t>(op, in, n, slots, out, out_f64, st);
        case CHAI_F32: return run_reduce_noalloc<float>(op, in, n, slots, out, out_f64, st);
        default: return run_reduce_noalloc<double>(op, in, n, slots, out, out_f64, st);
    }
}
