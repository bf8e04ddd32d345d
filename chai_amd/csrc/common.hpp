// common.hpp -- shared host-side helpers for the MI355X back end: status
// codes, per-thread error text, HIP error mapping, precision helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/chai_mi355x.h"

namespace chai_hip {

void set_error(const char* fmt, ...);

inline int hip_fail(hipError_t e, const char* what)
{
    set_error("%s: %s", what, hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? CHAI_ERR_NOMEM : CHAI_ERR_HIP;
}

#define CHAI_HIP_TRY(expr)                                               \
    do {                                                                 \
        hipError_t _e = (expr);                                          \
        if (_e != hipSuccess) return ::chai_hip::hip_fail(_e, #expr);    \
    } while (0)

// Launch check: reports launch-configuration errors without synchronising.
#define CHAI_LAUNCH_CHECK(name)                                          \
    do {                                                                 \
        hipError_t _e = hipGetLastError();                               \
        if (_e != hipSuccess) return ::chai_hip::hip_fail(_e, name);     \
    } while (0)

inline bool valid_prec(int p) { return p == CHAI_U32 || p == CHAI_I32 || p == CHAI_F32 || p == CHAI_F64; }
inline size_t prec_size(int p) { return p == CHAI_F64 ? 8 : 4; }
// PrecType::vecLength (misc/PrecType.cpp:29-48)
inline size_t std_vec_len(int p) { return p == CHAI_F64 ? 2 : 4; }
// PrecType::padWidth (misc/PrecType.cpp:120-133)
inline size_t pad_width(int p, size_t W)
{
    size_t v = std_vec_len(p);
    return (W % v) ? v * (1 + W / v) : W;
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned ceil_div_u(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace chai_hip
