// runtime.hip -- lifecycle, memory, timing entry points of the C ABI.
// Replaces vendor/OCLinit.cpp / vendor/OCLdevice.cpp buffer and queue
// handling and the gettimeofday timing of autotune/AutoTuneBench.cpp:116-123.
#include "common.hpp"

#include <cstring>
#include <mutex>

namespace chai_hip {

int reduce_scratch_setup(int device);   // reduce.hip

static thread_local char g_err[512];

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

}  // namespace chai_hip

using namespace chai_hip;

namespace {
std::mutex g_init_mu;
int g_device = -1;

struct Timer {
    hipEvent_t start = nullptr, stop = nullptr;
};
}  // namespace

extern "C" {

const char* chai_hip_last_error(void) { return g_err; }

int chai_hip_init(int device_ordinal)
{
    std::lock_guard<std::mutex> lk(g_init_mu);
    int n = 0;
    CHAI_HIP_TRY(hipGetDeviceCount(&n));
    if (device_ordinal < 0 || device_ordinal >= n) {
        set_error("chai_hip_init: device %d out of range (%d devices)", device_ordinal, n);
        return CHAI_ERR_ARG;
    }
    CHAI_HIP_TRY(hipSetDevice(device_ordinal));
    hipDeviceProp_t prop;
    CHAI_HIP_TRY(hipGetDeviceProperties(&prop, device_ordinal));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        set_error("chai_hip_init: device %d is %s; this library is built for gfx950 only",
                  device_ordinal, prop.gcnArchName);
        return CHAI_ERR_STATE;
    }
    // device address of the no-allocation reduction's scratch array
    int rc = reduce_scratch_setup(device_ordinal);
    if (rc != CHAI_OK) return rc;
    g_device = device_ordinal;
    return CHAI_OK;
}

int chai_hip_shutdown(void)
{
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (g_device >= 0) CHAI_HIP_TRY(hipDeviceSynchronize());
    g_device = -1;
    return CHAI_OK;
}

int chai_hip_device_info(int* cu_count, int* lds_per_cu, int* clock_khz, char* arch_name,
                         size_t arch_len)
{
    int dev = 0;
    CHAI_HIP_TRY(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    CHAI_HIP_TRY(hipGetDeviceProperties(&prop, dev));
    if (cu_count) *cu_count = prop.multiProcessorCount;
    if (lds_per_cu) *lds_per_cu = (int)prop.maxSharedMemoryPerMultiProcessor;
    if (clock_khz) *clock_khz = prop.clockRate;
    if (arch_name && arch_len) {
        std::strncpy(arch_name, prop.gcnArchName, arch_len - 1);
        arch_name[arch_len - 1] = 0;
    }
    return CHAI_OK;
}

int chai_hip_alloc(size_t bytes, void** out)
{
    if (!out) { set_error("chai_hip_alloc: null out"); return CHAI_ERR_ARG; }
    *out = nullptr;
    if (bytes == 0) return CHAI_OK;
    CHAI_HIP_TRY(hipMalloc(out, bytes));
    return CHAI_OK;
}

int chai_hip_free(void* ptr)
{
    if (ptr) CHAI_HIP_TRY(hipFree(ptr));
    return CHAI_OK;
}

int chai_hip_h2d(void* dst, const void* src, size_t bytes, void* stream)
{
    if (bytes == 0) return CHAI_OK;
    CHAI_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)));
    return CHAI_OK;
}

int chai_hip_d2h(void* dst, const void* src, size_t bytes, void* stream)
{
    if (bytes == 0) return CHAI_OK;
    CHAI_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
    CHAI_HIP_TRY(hipStreamSynchronize(as_stream(stream)));
    return CHAI_OK;
}

int chai_hip_memset(void* dst, int value, size_t bytes, void* stream)
{
    if (bytes == 0) return CHAI_OK;
    CHAI_HIP_TRY(hipMemsetAsync(dst, value, bytes, as_stream(stream)));
    return CHAI_OK;
}

int chai_hip_sync(void* stream)
{
    CHAI_HIP_TRY(hipStreamSynchronize(as_stream(stream)));
    return CHAI_OK;
}

int chai_hip_timer_create(void** timer)
{
    if (!timer) { set_error("chai_hip_timer_create: null out"); return CHAI_ERR_ARG; }
    Timer* t = new Timer;
    hipError_t e = hipEventCreate(&t->start);
    if (e == hipSuccess) e = hipEventCreate(&t->stop);
    if (e != hipSuccess) {
        if (t->start) (void)hipEventDestroy(t->start);
        delete t;
        return hip_fail(e, "hipEventCreate");
    }
    *timer = t;
    return CHAI_OK;
}

int chai_hip_timer_start(void* timer, void* stream)
{
    if (!timer) { set_error("timer_start: null timer"); return CHAI_ERR_ARG; }
    CHAI_HIP_TRY(hipEventRecord(static_cast<Timer*>(timer)->start, as_stream(stream)));
    return CHAI_OK;
}

int chai_hip_timer_stop(void* timer, void* stream)
{
    if (!timer) { set_error("timer_stop: null timer"); return CHAI_ERR_ARG; }
    CHAI_HIP_TRY(hipEventRecord(static_cast<Timer*>(timer)->stop, as_stream(stream)));
    return CHAI_OK;
}

int chai_hip_timer_elapsed_ms(void* timer, float* ms)
{
    if (!timer || !ms) { set_error("timer_elapsed: null argument"); return CHAI_ERR_ARG; }
    Timer* t = static_cast<Timer*>(timer);
    CHAI_HIP_TRY(hipEventSynchronize(t->stop));
    CHAI_HIP_TRY(hipEventElapsedTime(ms, t->start, t->stop));
    return CHAI_OK;
}

int chai_hip_timer_destroy(void* timer)
{
    if (!timer) return CHAI_OK;
    Timer* t = static_cast<Timer*>(timer);
    (void)hipEventDestroy(t->start);
    (void)hipEventDestroy(t->stop);
    delete t;
    return CHAI_OK;
}

}  // extern "C"
