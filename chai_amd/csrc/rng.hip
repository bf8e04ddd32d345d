// rng.hip -- chai_hip_rng_fill: materialise Chai's RNG stream into HBM.
//
// A thread generates one vector (4 x 32-bit or 2 x 64-bit elements) per step
// of a grid-stride loop, so a wavefront writes 64 contiguous 16-byte vectors =
// 1 KiB per store instruction when the row width is a multiple of the vector
// length (fully coalesced).
// vidx follows the width-padded layout of kernel/Subscript.cpp:359-380 so the
// stream is independent of how the kernel is tiled.
#include "common.hpp"
#include "rng_device.hpp"

#include <type_traits>

namespace chai_hip {

struct RngGeom {
    uint64_t W, H, slots;
    uint64_t vpr;        // vectors per (padded) row
    uint64_t nvec;       // total vectors = slots * H * vpr
    uint64_t seed;
    float lo_f, d_f;     // limits for f32
    double lo_d, d_d;    // limits for f64
    int has_limits, lo_nonzero;
};

// the elements of vector v of the stream.  LIM (f32 uniforms): 0 the unit
// interval, 1 scaled by d_f, 2 scaled and shifted by lo_f (a template
// argument, so the launch picks the arithmetic instead of every element)
template <typename T, int VARIANT, int VL, int NORMAL, int LIM>
__device__ __forceinline__ void rng_vals_32(const RngGeom& g, uint64_t v, T (&val)[4])
{
    uint32_t w[4];
    rng_words32<VARIANT, VL>((uint32_t)v, (uint32_t)g.seed, w);
    if constexpr (std::is_integral<T>::value) {
        for (int i = 0; i < 4; i++) val[i] = (T)w[i];
    } else if constexpr (NORMAL) {
        float f[4];
        normals_f32(w, VL == 4 ? 2 : 1, f);
        for (int i = 0; i < 4; i++) val[i] = f[i];
    } else {
        for (int i = 0; i < 4; i++) {
            float u = unit_f32(w[i]);
            if (LIM >= 1) u = u * g.d_f;
            if (LIM == 2) u = u + g.lo_f;
            val[i] = u;
        }
    }
}

// vector v of the stream -> its elements in `out`.  PAD: rows end inside a
// vector (W not a multiple of VL); without it the rows are contiguous in the
// vector index (element offset v * VL: no division, no row product)
template <typename T, int VARIANT, int VL, int NORMAL, int LIM = 0, bool PAD = true>
__device__ __forceinline__ void rng_vec_32(const RngGeom& g, T* __restrict__ out, uint64_t v)
{
    uint64_t row = 0, x0 = v * VL, n = VL;
    if (PAD) {
        row = v / g.vpr;                       // slot * H + y
        x0 = (v - row * g.vpr) * VL;
        if (x0 >= g.W) return;                 // a vector entirely in the padding
        n = (g.W - x0) < (uint64_t)VL ? (g.W - x0) : (uint64_t)VL;
    }
    T val[4];
    rng_vals_32<T, VARIANT, VL, NORMAL, LIM>(g, v, val);
    T* dst = PAD ? out + row * g.W + x0 : out + x0;
    if (VL == 4 && n == 4 && (g.W & 3) == 0 && ((uintptr_t)out & 15) == 0) {
        typedef T T4 __attribute__((ext_vector_type(4)));
        T4 q = {val[0], val[1], val[2], val[3]};
        *reinterpret_cast<T4*>(dst) = q;
    } else {
        for (uint64_t i = 0; i < n; i++) dst[i] = val[i];
    }
}

// One vector per thread (the launch covers the array), grid-stride past the
// grid's size.
template <typename T, int VARIANT, int VL, int NORMAL, int LIM>
__global__ __launch_bounds__(256) void rng_fill_32(RngGeom g, T* __restrict__ out)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t v0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g.W == g.vpr * VL) {
        for (uint64_t v = v0; v < g.nvec; v += stride) rng_vec_32<T, VARIANT, VL, NORMAL, LIM, false>(g, out, v);
    } else {
        for (uint64_t v = v0; v < g.nvec; v += stride) rng_vec_32<T, VARIANT, VL, NORMAL, LIM, true>(g, out, v);
    }
}

template <int VARIANT, int VL, int NORMAL>
__device__ __forceinline__ void rng_vec_64(const RngGeom& g, double* __restrict__ out, uint64_t v)
{
    uint64_t row = 0, x0 = v * VL, n = VL;
    if (g.W != g.vpr * VL) {
        row = v / g.vpr;
        x0 = (v - row * g.vpr) * VL;
        if (x0 >= g.W) return;
        n = (g.W - x0) < (uint64_t)VL ? (g.W - x0) : (uint64_t)VL;
    }
    uint64_t w[2];
    rng_words64<VARIANT>(v, g.seed, w);
    double val[2];
    const double u0 = unit_f64(w[0]), u1 = unit_f64(w[1]);
    if (NORMAL) {
        box_muller(u0, u1, val[0], val[1]);
    } else {
        val[0] = u0; val[1] = u1;
        if (g.has_limits) {
            for (int i = 0; i < 2; i++) {
                val[i] = val[i] * g.d_d;
                if (g.lo_nonzero) val[i] = val[i] + g.lo_d;
            }
        }
    }
    double* dst = out + row * g.W + x0;
    for (uint64_t i = 0; i < n; i++) dst[i] = val[i];
}

template <int VARIANT, int VL, int NORMAL>
__global__ __launch_bounds__(256) void rng_fill_64(RngGeom g, double* __restrict__ out)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < g.nvec; v += stride)
        rng_vec_64<VARIANT, VL, NORMAL>(g, out, v);
}

// blocks for n vectors: one vector per thread.  The waves resident at any
// time then write one compact stretch of the array; a grid of a few thousand
// blocks striding over it (every wave writing 1 KiB per step, 16 MiB apart
// between steps) stored 2^28 floats at 4.2-4.5 TB/s against 7.0 TB/s for this
// grid, with the Philox arithmetic alone taking 0.11 ms of the 0.15 ms.  A
// launch holds at most 2^31 work-items (the dispatch counts them in 32 bits);
// the kernels stride past that.
static unsigned rng_blocks(uint64_t nvec)
{
    const uint64_t b = (nvec + 255) / 256;
    return (unsigned)(b < (1ull << 23) ? b : (1ull << 23));
}

template <typename T, int VARIANT, int NORMAL, int LIM>
static void launch32_lim(int vl, const RngGeom& g, void* out, hipStream_t st)
{
    const unsigned blocks = rng_blocks(g.nvec);
    if (vl == 4) rng_fill_32<T, VARIANT, 4, NORMAL, LIM><<<blocks, 256, 0, st>>>(g, (T*)out);
    else if (vl == 2) rng_fill_32<T, VARIANT, 2, NORMAL, LIM><<<blocks, 256, 0, st>>>(g, (T*)out);
    else rng_fill_32<T, VARIANT, 1, NORMAL, LIM><<<blocks, 256, 0, st>>>(g, (T*)out);
}

template <typename T, int VARIANT, int NORMAL>
static void launch32(int vl, const RngGeom& g, void* out, hipStream_t st)
{
    if (std::is_floating_point<T>::value && !NORMAL && g.has_limits) {
        if (g.lo_nonzero) launch32_lim<T, VARIANT, NORMAL, 2>(vl, g, out, st);
        else launch32_lim<T, VARIANT, NORMAL, 1>(vl, g, out, st);
    } else {
        launch32_lim<T, VARIANT, NORMAL, 0>(vl, g, out, st);
    }
}

template <int VARIANT>
static int dispatch(int prec, int vl, int normal, const RngGeom& g, void* out, hipStream_t st)
{
    switch (prec) {
        case CHAI_U32: launch32<uint32_t, VARIANT, 0>(vl, g, out, st); break;
        case CHAI_I32: launch32<int32_t, VARIANT, 0>(vl, g, out, st); break;
        case CHAI_F32:
            if (normal) launch32<float, VARIANT, 1>(vl, g, out, st);
            else launch32<float, VARIANT, 0>(vl, g, out, st);
            break;
        case CHAI_F64: {
            const unsigned blocks = rng_blocks(g.nvec);
            if (vl == 2) {
                if (normal) rng_fill_64<VARIANT, 2, 1><<<blocks, 256, 0, st>>>(g, (double*)out);
                else rng_fill_64<VARIANT, 2, 0><<<blocks, 256, 0, st>>>(g, (double*)out);
            } else {
                if (normal) rng_fill_64<VARIANT, 1, 1><<<blocks, 256, 0, st>>>(g, (double*)out);
                else rng_fill_64<VARIANT, 1, 0><<<blocks, 256, 0, st>>>(g, (double*)out);
            }
            break;
        }
    }
    return CHAI_OK;
}

// Every 32-bit stream word through the f32 normals' log and sqrt and through
// the general forms (bm_logf with its IEEE divide, the correctly rounded
// sqrtf): cnt[0] counts the words whose logs differ in any bit, cnt[1] those
// whose square roots do.
__global__ __launch_bounds__(256) void bm_check_kernel(unsigned long long* cnt)
{
    const uint64_t nthr = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long bl = 0, bs = 0;
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < (1ull << 32); w += nthr) {
        const float u = unit_f32((uint32_t)w);
        const float l = bm_logf(u);
        bl += __float_as_uint(l) != __float_as_uint(bm_logf_unit(u));
        const float y = -2.0f * l;
        bs += __float_as_uint(__builtin_sqrtf(y)) != __float_as_uint(bm_sqrtf(y));
    }
    if (bl) atomicAdd(&cnt[0], bl);
    if (bs) atomicAdd(&cnt[1], bs);
}

}  // namespace chai_hip

using namespace chai_hip;

extern "C" int chai_hip_rng_check_normals(uint64_t* mismatches, void* stream)
{
    if (!mismatches) { set_error("rng_check_normals: null output"); return CHAI_ERR_ARG; }
    unsigned long long* d = nullptr;
    CHAI_HIP_TRY(hipMalloc(&d, 2 * sizeof(unsigned long long)));
    hipStream_t st = as_stream(stream);
    hipError_t e = hipMemsetAsync(d, 0, 2 * sizeof(unsigned long long), st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(bm_check_kernel, dim3(2048), dim3(256), 0, st, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(mismatches, d, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    const hipError_t ef = hipFree(d);
    if (e == hipSuccess) e = ef;
    if (e != hipSuccess) return hip_fail(e, "rng_check_normals");
    return CHAI_OK;
}

extern "C" int chai_hip_rng_fill(int variant, int prec, int vec_len, uint64_t seed, size_t W,
                                 size_t H, size_t slots, int normal, int has_limits, double lo,
                                 double hi, void* out, void* stream)
{
    if (variant != CHAI_RNG_PHILOX && variant != CHAI_RNG_THREEFRY) {
        set_error("rng_fill: unknown variant %d", variant);
        return CHAI_ERR_ARG;
    }
    if (!valid_prec(prec)) { set_error("rng_fill: bad precision %d", prec); return CHAI_ERR_ARG; }
    const int vl = vec_len ? vec_len : (int)std_vec_len(prec);
    const bool vl_ok = prec == CHAI_F64 ? (vl == 1 || vl == 2) : (vl == 1 || vl == 2 || vl == 4);
    if (!vl_ok) { set_error("rng_fill: vector length %d invalid for precision %d", vl, prec); return CHAI_ERR_ARG; }
    if (normal && prec != CHAI_F32 && prec != CHAI_F64) {
        set_error("rng_fill: normal distribution needs a floating point precision");
        return CHAI_ERR_ARG;
    }
    if (W == 0 || H == 0 || slots == 0) return CHAI_OK;
    if (!out) { set_error("rng_fill: null output"); return CHAI_ERR_ARG; }
    RngGeom g;
    g.W = W; g.H = H; g.slots = slots;
    g.vpr = pad_width(prec, W) / (size_t)vl;
    g.nvec = (uint64_t)slots * H * g.vpr;
    if (prec != CHAI_F64 && g.nvec > 0xffffffffull) {
        set_error("rng_fill: %llu vectors exceed the 32-bit vector index of the 32-bit stream",
                  (unsigned long long)g.nvec);
        return CHAI_ERR_ARG;
    }
    g.seed = seed;
    g.has_limits = has_limits && !normal;
    g.lo_nonzero = lo != 0.0;
    g.d_f = (float)(hi - lo); g.lo_f = (float)lo;
    g.d_d = hi - lo; g.lo_d = lo;
    hipStream_t st = as_stream(stream);
    if (variant == CHAI_RNG_PHILOX) dispatch<1230>(prec, vl, normal, g, out, st);
    else dispatch<1231>(prec, vl, normal, g, out, st);
    CHAI_LAUNCH_CHECK("rng_fill");
    return CHAI_OK;
}
