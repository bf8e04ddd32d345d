// trace_vm.hpp -- the fused-trace virtual machine (device side).
//
// One pre-compiled kernel executes any elementwise / reduction trace the host
// compiler (trace.hip) lowers into 3-address VM instructions.  Each thread
// owns a vector of 4 consecutive row elements (the reference JIT's float4
// vectors, enqueue/EnqueueIdSpace.cpp:19-320); all lanes of a wavefront run the
// same instruction, so dispatch is a uniform scalar branch and the instruction
// stream is read through the scalar cache.  Values live in 12 register slots
// R0..R11 (4 lanes x 64 bits each) addressed through uniform switches, so no
// per-thread array is ever indexed dynamically (nothing goes to scratch).
//
// Scalar semantics follow Chai's interpreter (misc/SimpleFun{1,2,3}.cpp):
// comparisons yield 0/1 in the operand type, integer shifts use the low five
// bits of the count (x86 behaviour of the reference build), saturating
// RHADD, non-rotating signed ROTATE, max/min as a<b?b:a.
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <cstdint>
#include "../../include/chai_mi355x.h"
#endif

#include "rng_device.hpp"
#include "reduce_device.hpp"

namespace chai_hip {

enum VmOp : uint16_t {
    VM_LOAD = 1,   // R[dst] <- var a (broadcast by modulo), element prec p
    VM_IMM,        // R[dst] <- immediate bits u
    VM_RNG,        // R[dst] <- rng stream: a = variant, b = vec_len, c = normal|limits<<1, rng dims index e
    VM_INDEX,      // R[dst] <- index along dim a, as prec p
    VM_CVT,        // R[dst] <- convert R[s1] from prec p2 to prec p (through double)
    VM_UN,         // R[dst] <- op a (R[s1])
    VM_BIN,        // R[dst] <- op a (R[s1], R[s2])
    VM_TER,        // R[dst] <- op a (R[s1], R[s2], R[s3])
    VM_SEL,        // R[dst] <- R[s1] != 0 ? R[s2] : R[s3]   (s1 in prec p2, values in p)
    VM_STORE,      // var a <- R[s1]
    VM_ACC,        // acc[a] += (double) R[s1]
    VM_MOV,        // R[dst] <- R[s1]
    VM_GATHER,     // R[dst] <- var a at subscripts R[s1] (prec b), R[s2] (prec c, e == 2); p2 = chain
};

struct VmIns {
    uint16_t op;
    uint8_t p, p2;
    uint8_t dst, s1, s2, s3;
    uint32_t a, b, c, e;
    uint64_t u;
    double d0, d1;
};

struct VmVar {
    const void* in;
    void* out;
    uint64_t W, H, S;
    uint32_t prec, same;       // same: dims equal the index space
};

struct VmRngDims {
    uint64_t W, H, S, vpr;     // vpr = padded width / vec_len
    uint32_t same, pad;
};

// Nested broadcast: the interpreter evaluates every operator into an array of
// its own dims (the max of its operands') and reads operands by modulo
// (interp/BaseInterp.cpp:50-55), so a leaf under intermediate results smaller
// than the index space sees the coordinate x % d1 % d2 % ... (one modulus per
// smaller ancestor, outermost first).  A chain holds those moduli per
// dimension, already reduced (a modulus divisible by the next one is dropped).
struct VmChain {
    uint64_t W[3], H[3], S[3];
    uint32_t nW, nH, nS, pad;
};

__device__ __forceinline__ uint64_t chain_mod(uint64_t v, const uint64_t* m, uint32_t n)
{
#pragma unroll
    for (int l = 0; l < 3; l++)
        if (l < (int)n) v %= m[l];
    return v;
}

constexpr int VM_MAX_ACC = 4;
constexpr int VM_THREADS = 256;

struct VmLaunch {
    const VmIns* prog;
    uint32_t nins, nacc;
    const VmVar* vars;
    const VmRngDims* rng;
    uint64_t W, H, S, vpr, vps;         // index space; vectors per row / per slot
    double* partial;                    // [S][gridDim.x][nacc]
};

struct V4 {
    uint64_t l[4];
};

// ------------------------------------------------------------ bit helpers
__device__ __forceinline__ float f_(uint64_t b) { return __uint_as_float((uint32_t)b); }
__device__ __forceinline__ double d_(uint64_t b) { return __longlong_as_double((long long)b); }
__device__ __forceinline__ int32_t i_(uint64_t b) { return (int32_t)(uint32_t)b; }
__device__ __forceinline__ uint32_t u_(uint64_t b) { return (uint32_t)b; }
__device__ __forceinline__ uint64_t bf(float v) { return (uint64_t)__float_as_uint(v); }
__device__ __forceinline__ uint64_t bd(double v) { return (uint64_t)__double_as_longlong(v); }
__device__ __forceinline__ uint64_t bi(int32_t v) { return (uint64_t)(uint32_t)v; }
__device__ __forceinline__ uint64_t bu(uint32_t v) { return (uint64_t)v; }

__device__ __forceinline__ double as_double(int prec, uint64_t b)
{
    switch (prec) {
        case 0: return (double)u_(b);
        case 1: return (double)i_(b);
        case 4: return (double)f_(b);
        default: return d_(b);
    }
}

// Float -> integer conversions as the interpreter's x86 host performs them
// (cvttsd2si / cvttss2si): NaN and out-of-range values give the "integer
// indefinite" (the most negative value) instead of the GPU's clamped result.
__device__ __forceinline__ long long x86_ll(double v)
{
    return (v != v || v >= 0x1p63 || v < -0x1p63) ? (long long)0x8000000000000000ull : (long long)v;
}
__device__ __forceinline__ int x86_i(double v)
{
    return (v != v || v >= 0x1p31 || v < -0x1p31) ? (int)0x80000000u : (int)v;
}

// double -> element bits, C conversion semantics of the interpreter
// (InterpConvert.cpp: values pass through double)
__device__ __forceinline__ uint64_t from_double(int prec, double v)
{
    switch (prec) {
        case 0: return bu((uint32_t)x86_ll(v));
        case 1: return bi((int32_t)x86_ll(v));
        case 4: return bf((float)v);
        default: return bd(v);
    }
}

// ---------------------------------------------------------- scalar ops
__device__ __forceinline__ uint32_t shl32(uint32_t a, uint32_t n) { return a << (n & 31u); }
__device__ __forceinline__ uint32_t shr32u(uint32_t a, uint32_t n) { return a >> (n & 31u); }
__device__ __forceinline__ int32_t shr32s(int32_t a, uint32_t n) { return a >> (n & 31u); }

// Ops whose device math is large (transcendentals other than exp / log and
// the remaining SimpleFun1 functions) are kept out of line in the VM kernel,
// whose generic dispatch would otherwise inline every one of them; the JIT
// (one op per call site, known at compile time) inlines everything.
#ifdef __HIPCC_RTC__
#define CHAI_VM_HEAVY __forceinline__
#else
#define CHAI_VM_HEAVY __noinline__
#endif

// pown in double: square-and-multiply from the low bit, reciprocal for n < 0
// (what the reference's pow(double, int) resolves to in its C++98 build)
__device__ __forceinline__ double powi_sq(double x, int m)
{
    unsigned n = m < 0 ? 0u - (unsigned)m : (unsigned)m;
    double y = (n & 1u) ? x : 1.0;
    while (n >>= 1) {
        x = x * x;
        if (n & 1u) y = y * x;
    }
    return m < 0 ? 1.0 / y : y;
}

// pi and 1/pi as the reference's M_PI / M_1_PI (double)
constexpr double VM_PI = 0x1.921fb54442d18p+1;
constexpr double VM_1_PI = 0x1.45f306dc9c883p-2;

template <typename T>
__device__ CHAI_VM_HEAVY T fp_bin_heavy(uint32_t op, T a, T b)
{
    switch (op) {
        case 248: return sizeof(T) == 4 ? (T)powf((float)a, (float)b) : (T)pow((double)a, (double)b);
        case 207: return sizeof(T) == 4 ? (T)atan2f((float)a, (float)b) : (T)atan2((double)a, (double)b);
        case 229: return sizeof(T) == 4 ? (T)fmodf((float)a, (float)b) : (T)fmod((double)a, (double)b);
        case 232: return sizeof(T) == 4 ? (T)hypotf((float)a, (float)b) : (T)hypot((double)a, (double)b);
        case 210: return sizeof(T) == 4 ? (T)((double)atan2f((float)a, (float)b) / VM_PI)
                                        : (T)(atan2((double)a, (double)b) / VM_PI);
        case 234: return sizeof(T) == 4 ? (T)ldexpf((float)a, x86_i(b)) : (T)ldexp((double)a, x86_i(b));
        case 247: return sizeof(T) == 4 ? (T)nextafterf((float)a, (float)b) : (T)nextafter((double)a, (double)b);
        case 249: return sizeof(T) == 4 ? (T)powf((float)a, (float)x86_i(b)) : (T)powi_sq((double)a, x86_i(b));
        case 250: return sizeof(T) == 4 ? (T)powf((float)a, (float)b) : (T)pow((double)a, (double)b);
        case 251: return sizeof(T) == 4 ? (T)remainderf((float)a, (float)b) : (T)remainder((double)a, (double)b);
        case 254: return sizeof(T) == 4 ? (T)powf((float)a, 1.0f / (float)b) : (T)pow((double)a, 1.0 / (double)b);
        default: return (T)0;
    }
}

__device__ __forceinline__ bool bin_is_heavy(uint32_t op)
{
    return op == 248 || op == 207 || op == 229 || op == 232 || op == 210 || op == 234 || op == 247 ||
           op == 249 || op == 250 || op == 251 || op == 254;
}

template <typename T>
__device__ __forceinline__ T fp_bin(uint32_t op, T a, T b)
{
    if (bin_is_heavy(op)) return fp_bin_heavy<T>(op, a, b);
    switch (op) {
        case 11: return a + b;
        case 12: return a / b;
        case 13: return a * b;
        case 14: return a - b;
        case 32: return (T)(a == b);
        case 33: return (T)(a != b);
        case 34: return (T)(a >= b);
        case 35: return (T)(a > b);
        case 36: return (T)(a <= b);
        case 37: return (T)(a < b);
        case 302: return a < b ? b : a;
        case 303: return a < b ? a : b;
        case 227: return sizeof(T) == 4 ? (T)fmaxf((float)a, (float)b) : (T)fmax((double)a, (double)b);
        case 228: return sizeof(T) == 4 ? (T)fminf((float)a, (float)b) : (T)fmin((double)a, (double)b);
        case 213: return sizeof(T) == 4 ? (T)copysignf((float)a, (float)b) : (T)copysign((double)a, (double)b);
        case 224: return sizeof(T) == 4 ? (T)fdimf((float)a, (float)b) : (T)fdim((double)a, (double)b);
        // misc/SimpleFun2.cpp definitions (the pi variant divides in double)
        case 243: {
            const T fa = a < 0 ? -a : a, fb = b < 0 ? -b : b;
            if (fa > fb) return a;
            if (fa < fb) return b;
            return sizeof(T) == 4 ? (T)fmaxf((float)a, (float)b) : (T)fmax((double)a, (double)b);
        }
        case 244: {
            const T fa = a < 0 ? -a : a, fb = b < 0 ? -b : b;
            if (fa < fb) return a;
            if (fa > fb) return b;
            return sizeof(T) == 4 ? (T)fminf((float)a, (float)b) : (T)fmin((double)a, (double)b);
        }
        case 306: return b < a ? (T)0 : (T)1;
        case 500: return (T)(a == b);
        case 501: return (T)(a != b);
        case 502: return (T)(a > b);
        case 503: return (T)(a >= b);
        case 504: return (T)(a < b);
        case 505: return (T)(a <= b);
        case 506: return (T)((a < b) || (a > b));
        case 511: return (T)(a == a && b == b);
        case 512: return (T)(a != a || b != b);
        default: return (T)0;
    }
}

__device__ __forceinline__ uint32_t u_bin(uint32_t op, uint32_t a, uint32_t b)
{
    switch (op) {
        case 11: return a + b;
        case 12: return b ? a / b : 0u;
        case 13: return a * b;
        case 14: return a - b;
        case 17: return b ? a % b : 0u;
        case 18: return shl32(a, b);
        case 19: return shr32u(a, b);
        case 20: return a & b;
        case 21: return a | b;
        case 22: return a ^ b;
        case 30: return (uint32_t)(a && b);
        case 31: return (uint32_t)(a || b);
        case 32: return (uint32_t)(a == b);
        case 33: return (uint32_t)(a != b);
        case 34: return (uint32_t)(a >= b);
        case 35: return (uint32_t)(a > b);
        case 36: return (uint32_t)(a <= b);
        case 37: return (uint32_t)(a < b);
        case 302: return a < b ? b : a;
        case 303: return b < a ? b : a;
        case 401: return a > b ? a - b : b - a;
        case 402: { const uint64_t c = (uint64_t)a + b; return c > 0xffffffffull ? 0xffffffffu : (uint32_t)c; }
        case 405: { const uint64_t c = (uint64_t)a + b; return (c <= 0xffffffffull ? (uint32_t)c : 0xffffffffu) >> 1; }
        case 411: return (a & 0x00ffffffu) * (b & 0x00ffffffu);
        case 412: return (uint32_t)(((uint64_t)a * b) >> 32);
        case 413: { const uint64_t c = (uint64_t)a + b + 1; return (uint32_t)((c <= 0xffffffffull ? c : 0xffffffffull) >> 1); }
        case 414: return shl32(a, b) | shr32u(a, 32u - b);
        case 415: return a < b ? 0u : b - a;   // reference returns b - a (SimpleFun2.cpp:521-527)
        default: return 0u;
    }
}

__device__ __forceinline__ int32_t i_bin(uint32_t op, int32_t a, int32_t b)
{
    const uint32_t ua = (uint32_t)a, ub = (uint32_t)b;
    switch (op) {
        case 11: return (int32_t)(ua + ub);
        case 12: return (b == 0 || (a == INT32_MIN && b == -1)) ? 0 : a / b;
        case 13: return (int32_t)(ua * ub);
        case 14: return (int32_t)(ua - ub);
        case 17: return (b == 0 || (a == INT32_MIN && b == -1)) ? 0 : a % b;
        case 18: return (int32_t)shl32(ua, ub);
        case 19: return shr32s(a, ub);
        case 20: return a & b;
        case 21: return a | b;
        case 22: return a ^ b;
        case 30: return (int32_t)(a && b);
        case 31: return (int32_t)(a || b);
        case 32: return (int32_t)(a == b);
        case 33: return (int32_t)(a != b);
        case 34: return (int32_t)(a >= b);
        case 35: return (int32_t)(a > b);
        case 36: return (int32_t)(a <= b);
        case 37: return (int32_t)(a < b);
        case 302: return a < b ? b : a;
        case 303: return b < a ? b : a;
        case 401: return (int32_t)(a > b ? ua - ub : ub - ua);
        // upper saturation only; below INT32_MIN the 32-bit value wraps
        case 402: { const long long c = (long long)a + b; return c <= INT32_MAX ? (int32_t)(uint32_t)c : INT32_MAX; }
        case 405: { const long long c = (long long)a + b; const int32_t d = c <= INT32_MAX ? (int32_t)(uint32_t)c : INT32_MAX; return d >> 1; }
        case 411: return (int32_t)((uint32_t)(a & 0x00ffffff) * (uint32_t)(b & 0x00ffffff));
        case 412: return (int32_t)(((long long)a * b) >> 32);
        case 413: { const long long c = (long long)a + b + 1; const int32_t d = c <= INT32_MAX ? (int32_t)(uint32_t)c : INT32_MAX; return d >> 1; }
        case 414: return (int32_t)shl32(ua, ub) | shr32s(a, 32u - ub);
        case 415: { const long long c = (long long)a - b; return c > INT32_MAX ? INT32_MAX : c < INT32_MIN ? INT32_MIN : (int32_t)c; }
        default: return 0;
    }
}

__device__ __forceinline__ bool un_is_light(uint32_t op)
{
    return op == 10 || op == 223 || op == 261 || op == 256 || op == 219 || op == 220 || op == 237 ||
           op == 238 || op == 225 || op == 212 || op == 253 || op == 255 || op == 266;
}

template <typename T>
__device__ CHAI_VM_HEAVY T fp_un_heavy(uint32_t op, T a)
{
    const bool f = sizeof(T) == 4;
    switch (op) {
        case 221: return f ? (T)exp10f((float)a) : (T)exp10((double)a);
        case 222: return f ? (T)expm1f((float)a) : (T)expm1((double)a);
        case 239: return f ? (T)log10f((float)a) : (T)log10((double)a);
        case 240: return f ? (T)log1pf((float)a) : (T)log1p((double)a);
        case 257: return f ? (T)sinf((float)a) : (T)sin((double)a);
        case 214: return f ? (T)cosf((float)a) : (T)cos((double)a);
        case 262: return f ? (T)tanf((float)a) : (T)tan((double)a);
        case 203: return f ? (T)asinf((float)a) : (T)asin((double)a);
        case 200: return f ? (T)acosf((float)a) : (T)acos((double)a);
        case 206: return f ? (T)atanf((float)a) : (T)atan((double)a);
        case 259: return f ? (T)sinhf((float)a) : (T)sinh((double)a);
        case 215: return f ? (T)coshf((float)a) : (T)cosh((double)a);
        case 263: return f ? (T)tanhf((float)a) : (T)tanh((double)a);
        case 211: return f ? (T)cbrtf((float)a) : (T)cbrt((double)a);
        case 218: return f ? (T)erff((float)a) : (T)erf((double)a);
        case 217: return f ? (T)erfcf((float)a) : (T)erfc((double)a);
        // misc/SimpleFun1.cpp definitions (pi variants, degrees, radians in double)
        case 201: return f ? (T)acoshf((float)a) : (T)acosh((double)a);
        case 202: return f ? (T)((double)acosf((float)a) / VM_PI) : (T)(acos((double)a) / VM_PI);
        case 204: return f ? (T)asinhf((float)a) : (T)asinh((double)a);
        case 205: return f ? (T)((double)asinf((float)a) / VM_PI) : (T)(asin((double)a) / VM_PI);
        case 208: return f ? (T)atanhf((float)a) : (T)atanh((double)a);
        case 209: return f ? (T)((double)atanf((float)a) / VM_PI) : (T)(atan((double)a) / VM_PI);
        case 216: return f ? (T)cosf((float)(VM_PI * (double)a)) : (T)cos(VM_PI * (double)a);
        // ilogb: the host library's FP_ILOGB0 / FP_ILOGBNAN (INT_MIN) and INT_MAX for inf
        case 233: return a != a || a == 0 ? (T)-2147483648.0 : __builtin_isinf(a) ? (T)2147483647.0
                                          : f ? (T)ilogbf((float)a) : (T)ilogb((double)a);
        case 235: return f ? (T)lgammaf((float)a) : (T)lgamma((double)a);
        case 241: return f ? (T)logbf((float)a) : (T)logb((double)a);
        case 246: return (T)__builtin_nan("");
        case 260: return f ? (T)sinf((float)(VM_PI * (double)a)) : (T)sin(VM_PI * (double)a);
        case 264: return f ? (T)tanf((float)(VM_PI * (double)a)) : (T)tan(VM_PI * (double)a);
        case 265: return f ? (T)tgammaf((float)a) : (T)tgamma((double)a);
        case 301: return (T)(VM_1_PI * 180 * (double)a);
        case 305: return (T)((double)a * VM_PI / 180);
        case 308: return a > 0 ? (T)1 : a < 0 ? (T)-1 : a;
        case 507: return (T)(__builtin_isfinite(a) ? 1 : 0);
        case 508: return (T)(__builtin_isinf(a) ? 1 : 0);
        case 509: return (T)(a != a ? 1 : 0);
        case 510: return (T)(__builtin_isnormal(a) ? 1 : 0);
        case 513: return __builtin_signbit(a) ? (sizeof(T) == 4 ? (T)-2147483648.0f : (T)1) : (T)0;
        default: return (T)0;
    }
}

template <typename T>
__device__ __forceinline__ T fp_un(uint32_t op, T a)
{
    if (!un_is_light(op)) return fp_un_heavy<T>(op, a);
    const bool f = sizeof(T) == 4;
    switch (op) {
        case 10: return -a;
        case 223: return f ? (T)fabsf((float)a) : (T)fabs((double)a);
        case 261: return f ? (T)sqrtf((float)a) : (T)sqrt((double)a);
        case 256: return f ? (T)(1.0f / sqrtf((float)a)) : (T)(1.0 / sqrt((double)a));
        case 219: return f ? (T)expf((float)a) : (T)exp((double)a);
        case 220: return f ? (T)exp2f((float)a) : (T)exp2((double)a);
        case 237: return f ? (T)logf((float)a) : (T)log((double)a);
        case 238: return f ? (T)log2f((float)a) : (T)log2((double)a);
        case 225: return f ? (T)floorf((float)a) : (T)floor((double)a);
        case 212: return f ? (T)ceilf((float)a) : (T)ceil((double)a);
        case 253: return f ? (T)rintf((float)a) : (T)rint((double)a);
        case 255: return f ? (T)roundf((float)a) : (T)round((double)a);
        case 266: return f ? (T)truncf((float)a) : (T)trunc((double)a);
        default: return (T)0;
    }
}

__device__ __forceinline__ uint32_t u_un(uint32_t op, uint32_t a)
{
    switch (op) {
        case 10: return 0u - a;
        case 15: return ~a;
        case 16: return (uint32_t)(!a);
        case 400: return a;
        case 404: return a ? (uint32_t)__clz((int)a) : 32u;
        case 225: return a;
        default: return 0u;
    }
}

__device__ __forceinline__ int32_t i_un(uint32_t op, int32_t a)
{
    switch (op) {
        case 10: return (int32_t)(0u - (uint32_t)a);
        case 15: return ~a;
        case 16: return (int32_t)(!a);
        case 400: return a < 0 ? (int32_t)(0u - (uint32_t)a) : a;
        case 404: return a ? __clz(a) : 32;
        case 225: return a;
        default: return 0;
    }
}

template <typename T>
__device__ __forceinline__ T ter(uint32_t op, T a, T b, T c)
{
    switch (op) {
        case 300: { const T m = a < b ? b : a; return m < c ? m : c; }
        case 226:
            if constexpr (sizeof(T) == 4 && (T)0.5 != 0) return (T)__builtin_fmaf((float)a, (float)b, (float)c);
            else if constexpr ((T)0.5 != 0) return (T)__builtin_fma((double)a, (double)b, (double)c);
            else return (T)0;
        case 242: return a * b + c;
        case 304: { const T d = b - a; const T e = d * c; return a + e; }   // mix
        case 307: {                                                           // smoothstep(e0, e1, x)
            const T q = (c - a) / (b - a);
            const T q0 = q < 0 ? (T)0 : q, t = q0 < 1 ? q0 : (T)1;
            const T t2 = t * t, u = (T)3 - (T)2 * t;
            return t2 * u;
        }
        default: return (T)0;
    }
}

// integer ternary ops (misc/SimpleFun3.cpp): clamp, mad24, mad_hi, mad_sat
__device__ __forceinline__ uint32_t u_ter(uint32_t op, uint32_t a, uint32_t b, uint32_t c)
{
    switch (op) {
        case 300: { const uint32_t m = a < b ? b : a; return m < c ? m : c; }
        case 406: return u_bin(411, a, b) + c;
        case 407: return u_bin(412, a, b) + c;
        case 408: { const uint64_t d = (uint64_t)a * b + c; return d > 0xffffffffull ? 0xffffffffu : (uint32_t)d; }
        default: return 0u;
    }
}

__device__ __forceinline__ int32_t i_ter(uint32_t op, int32_t a, int32_t b, int32_t c)
{
    switch (op) {
        case 300: { const int32_t m = a < b ? b : a; return m < c ? m : c; }
        case 406: return (int32_t)((uint32_t)i_bin(411, a, b) + (uint32_t)c);
        case 407: return (int32_t)((uint32_t)i_bin(412, a, b) + (uint32_t)c);
        case 408: {
            const long long d = (long long)a * b + c;
            return d < INT32_MIN ? INT32_MIN : d > INT32_MAX ? INT32_MAX : (int32_t)d;
        }
        default: return 0;
    }
}

// ------------------------------------------------ loads, stores, rng, ops
__device__ __forceinline__ uint64_t vm_load_elem(const void* p, uint32_t prec, uint64_t idx)
{
    if (prec == CHAI_F64) return ((const uint64_t*)p)[idx];
    return (uint64_t)((const uint32_t*)p)[idx];
}

// 32-bit stream element from the words of its vector (component comp)
__device__ __forceinline__ uint64_t rng_elem32(const uint32_t w[4], int comp, int prec, int normal,
                                               int has_lim, double lo, double hi)
{
    // select instead of indexing with a runtime component (keeps w in registers)
    const uint32_t wc = comp == 0 ? w[0] : comp == 1 ? w[1] : comp == 2 ? w[2] : w[3];
    if (prec == CHAI_U32 || prec == CHAI_I32) return (uint64_t)wc;
    if (normal) {
        const uint32_t wa = comp < 2 ? w[0] : w[2], wb = comp < 2 ? w[1] : w[3];
        float n0, n1;
        box_muller_f32(unit_f32(wa), unit_f32(wb), n0, n1);
        return bf((comp & 1) ? n1 : n0);
    }
    float u = unit_f32(wc);
    if (has_lim) {
        u = u * (float)(hi - lo);
        if (lo != 0.0) u = u + (float)lo;
    }
    return bf(u);
}

__device__ __forceinline__ uint64_t rng_elem64(const uint64_t w[2], int comp, int normal, int has_lim,
                                               double lo, double hi)
{
    const double u0 = unit_f64(w[0]), u1 = unit_f64(w[1]);
    if (normal) {
        double n0, n1;
        box_muller(u0, u1, n0, n1);
        return bd(comp ? n1 : n0);
    }
    double u = comp ? u1 : u0;
    if (has_lim) {
        u = u * (hi - lo);
        if (lo != 0.0) u = u + lo;
    }
    return bd(u);
}

__device__ __forceinline__ void words32(int variant, int vl, uint32_t vidx, uint32_t seed, uint32_t w[4])
{
    if (variant == CHAI_RNG_THREEFRY) {
        if (vl == 4) rng_words32<1231, 4>(vidx, seed, w);
        else rng_words32<1231, 2>(vidx, seed, w);
    } else {
        if (vl == 4) rng_words32<1230, 4>(vidx, seed, w);
        else rng_words32<1230, 2>(vidx, seed, w);
    }
}

__device__ __forceinline__ void words64(int variant, uint64_t vidx, uint64_t seed, uint64_t w[2])
{
    if (variant == CHAI_RNG_THREEFRY) rng_words64<1231>(vidx, seed, w);
    else rng_words64<1230>(vidx, seed, w);
}

// Instruction fields arrive by value so the instruction never needs an
// address (an addressed copy would live in scratch and lose its uniformity).
struct RngArgs {
    uint32_t variant, vl, flags, prec;
    uint64_t seed;
    double lo, hi;
};

__device__ __forceinline__ V4 vm_rng(const RngArgs A, const VmRngDims R, uint64_t slot, uint64_t y,
                                     uint64_t x0, int nval, const VmChain C)
{
    V4 r = {{0, 0, 0, 0}};
    const bool chained = (C.nW | C.nH | C.nS) != 0;
    const int variant = (int)A.variant, vl = (int)A.vl, normal = A.flags & 1, lim = (A.flags >> 1) & 1;
    const int prec = (int)A.prec;
    struct { uint64_t u; double d0, d1; } I = {A.seed, A.lo, A.hi};
    if (prec != CHAI_F64) {
        if (R.same && vl == 4 && !chained) {   // one stream vector covers the thread's 4 lanes
            uint32_t w[4];
            words32(variant, 4, (uint32_t)(slot * R.H * R.vpr + y * R.vpr + x0 / 4), (uint32_t)I.u, w);
            if (normal && prec == CHAI_F32) {
                float n0, n1, n2, n3;
                box_muller_f32(unit_f32(w[0]), unit_f32(w[1]), n0, n1);
                box_muller_f32(unit_f32(w[2]), unit_f32(w[3]), n2, n3);
                r.l[0] = bf(n0); r.l[1] = bf(n1); r.l[2] = bf(n2); r.l[3] = bf(n3);
            } else {
#pragma unroll
                for (int i = 0; i < 4; i++) r.l[i] = rng_elem32(w, i, prec, 0, lim, I.d0, I.d1);
            }
            return r;
        }
        for (int i = 0; i < 4; i++) if (i < nval) {
            const uint64_t xr = chain_mod(x0 + i, C.W, C.nW) % R.W;
            const uint64_t yr = chain_mod(y, C.H, C.nH) % R.H, sr = chain_mod(slot, C.S, C.nS) % R.S;
            uint32_t w[4];
            words32(variant, vl, (uint32_t)(sr * R.H * R.vpr + yr * R.vpr + xr / vl), (uint32_t)I.u, w);
            r.l[i] = rng_elem32(w, (int)(xr % vl), prec, normal, lim, I.d0, I.d1);
        }
        return r;
    }
    for (int i = 0; i < 4; i++) if (i < nval) {
        const uint64_t xr = chain_mod(x0 + i, C.W, C.nW) % R.W;
        const uint64_t yr = chain_mod(y, C.H, C.nH) % R.H, sr = chain_mod(slot, C.S, C.nS) % R.S;
        uint64_t w[2];
        words64(variant, sr * R.H * R.vpr + yr * R.vpr + xr / vl, I.u, w);
        r.l[i] = rng_elem64(w, (int)(xr % vl), normal, lim, I.d0, I.d1);
    }
    return r;
}

// load through a broadcast chain (never the `same` case: a leaf with the dims
// of the index space has no smaller ancestor)
__device__ __forceinline__ V4 vm_load_c(const VmVar& V, const VmChain& C, uint64_t slot, uint64_t y,
                                        uint64_t x0, int nval)
{
    V4 r = {{0, 0, 0, 0}};
    const uint64_t sr = chain_mod(slot, C.S, C.nS) % V.S, yr = chain_mod(y, C.H, C.nH) % V.H;
    const uint64_t rowoff = (sr * V.H + yr) * V.W;
    for (int i = 0; i < 4; i++)
        if (i < nval) r.l[i] = vm_load_elem(V.in, V.prec, rowoff + chain_mod(x0 + i, C.W, C.nW) % V.W);
    return r;
}

__device__ __forceinline__ V4 vm_load(const VmVar& V, uint64_t slot, uint64_t y, uint64_t x0, int nval,
                                      uint64_t W, uint64_t H)
{
    V4 r = {{0, 0, 0, 0}};
    if (V.same) {
        const uint64_t base = (slot * H + y) * W + x0;
        if (nval == 4 && V.prec != CHAI_F64 && ((W & 3) == 0) && (((uintptr_t)V.in & 15) == 0)) {
            typedef uint32_t u4 __attribute__((ext_vector_type(4)));
            const u4 q = __builtin_nontemporal_load(reinterpret_cast<const u4*>((const uint32_t*)V.in + base));
            r.l[0] = q[0]; r.l[1] = q[1]; r.l[2] = q[2]; r.l[3] = q[3];
            return r;
        }
        for (int i = 0; i < 4; i++) if (i < nval) r.l[i] = vm_load_elem(V.in, V.prec, base + i);
        return r;
    }
    if (V.W == 1 && V.H == 1 && V.S == 1) {
        const uint64_t e = vm_load_elem(V.in, V.prec, 0);
        r.l[0] = r.l[1] = r.l[2] = r.l[3] = e;
        return r;
    }
    const uint64_t rowoff = ((slot % V.S) * V.H + (y % V.H)) * V.W;
    for (int i = 0; i < 4; i++) if (i < nval) r.l[i] = vm_load_elem(V.in, V.prec, rowoff + (x0 + i) % V.W);
    return r;
}

// Gather subscript (interp/InterpGather.hpp:24-26, misc/SimpleFun1.cpp:196-216):
// floor the element (identity for integers) and convert it to int as the
// reference's x86 build does (NaN and out-of-range values give INT_MIN, uint32
// reinterprets its bits); then "xf < 0 ? xf + W : xf", evaluated in size_t and
// stored to an int, and "% W" on that int sign-extended to 64 bits.  The result
// is always below W.
__device__ __forceinline__ uint64_t gather_sub(uint32_t prec, uint64_t b, uint64_t W)
{
    int xf;
    if (prec == CHAI_U32 || prec == CHAI_I32) xf = i_(b);
    else if (prec == CHAI_F32) xf = x86_i((double)floorf(f_(b)));
    else xf = x86_i(floor(d_(b)));
    if (xf >= 0) return (uint64_t)xf < W ? (uint64_t)xf : (uint64_t)xf % W;
    const int m = (int)(uint32_t)((uint64_t)(int64_t)xf + W);
    return (uint64_t)(int64_t)m % W;
}

// gather1_floor / gather2_floor: the output has the data array's shape; C
// reduces the caller's (y, slot) to the gather's own coordinate
__device__ __forceinline__ V4 vm_gather(const VmVar& V, const VmChain& C, uint64_t slot, uint64_t y, int two,
                                        uint32_t px, const V4& ix, uint32_t py, const V4& iy, int nval)
{
    V4 r = {{0, 0, 0, 0}};
    const uint64_t sr = chain_mod(slot, C.S, C.nS) % V.S, yr = chain_mod(y, C.H, C.nH) % V.H;
    uint64_t e[4];
    for (int i = 0; i < 4; i++) {
        const uint64_t xs = gather_sub(px, ix.l[i], V.W);
        const uint64_t ys = two ? gather_sub(py, iy.l[i], V.H) : yr;
        e[i] = (sr * V.H + ys) * V.W + xs;
    }
    // four consecutive elements (a shifted stencil read): one 16-byte load
    // (two for f64) at element alignment instead of four scalar loads
    if (nval == 4 && e[1] == e[0] + 1 && e[2] == e[0] + 2 && e[3] == e[0] + 3) {
        if (V.prec == CHAI_F64) {
            typedef uint64_t u2a __attribute__((ext_vector_type(2), aligned(8)));
            const u2a* q = reinterpret_cast<const u2a*>((const uint64_t*)V.in + e[0]);
            const u2a lo = q[0], hi = q[1];
            r.l[0] = lo[0]; r.l[1] = lo[1]; r.l[2] = hi[0]; r.l[3] = hi[1];
        } else {
            typedef uint32_t u4a __attribute__((ext_vector_type(4), aligned(4)));
            const u4a q = *reinterpret_cast<const u4a*>((const uint32_t*)V.in + e[0]);
            r.l[0] = q[0]; r.l[1] = q[1]; r.l[2] = q[2]; r.l[3] = q[3];
        }
        return r;
    }
    for (int i = 0; i < 4; i++)
        if (i < nval) r.l[i] = vm_load_elem(V.in, V.prec, e[i]);
    return r;
}

__device__ __forceinline__ void vm_store(const VmVar& V, const V4& v, uint64_t slot, uint64_t y,
                                         uint64_t x0, int nval, uint64_t W, uint64_t H)
{
    const uint64_t base = (slot * H + y) * W + x0;
    if (V.prec == CHAI_F64) {
        for (int i = 0; i < 4; i++) if (i < nval) ((uint64_t*)V.out)[base + i] = v.l[i];
    } else if (nval == 4 && (W & 3) == 0 && (((uintptr_t)V.out & 15) == 0)) {
        typedef uint32_t u4 __attribute__((ext_vector_type(4)));
        u4 q = {(uint32_t)v.l[0], (uint32_t)v.l[1], (uint32_t)v.l[2], (uint32_t)v.l[3]};
        __builtin_nontemporal_store(q, reinterpret_cast<u4*>((uint32_t*)V.out + base));
    } else {
        for (int i = 0; i < 4; i++) if (i < nval) ((uint32_t*)V.out)[base + i] = (uint32_t)v.l[i];
    }
}

__device__ __forceinline__ V4 vm_bin(uint32_t op, int prec, const V4& a, const V4& b)
{
    V4 r;
    switch (prec) {
        case CHAI_F32:
#pragma unroll
            for (int i = 0; i < 4; i++) r.l[i] = bf(fp_bin<float>(op, f_(a.l[i]), f_(b.l[i])));
            break;
        case CHAI_F64:
#pragma unroll
            for (int i = 0; i < 4; i++) r.l[i] = bd(fp_bin<double>(op, d_(a.l[i]), d_(b.l[i])));
            break;
        case CHAI_I32:
#pragma unroll
            for (int i = 0; i < 4; i++) r.l[i] = bi(i_bin(op, i_(a.l[i]), i_(b.l[i])));
            break;
        default:
#pragma unroll
            for (int i = 0; i < 4; i++) r.l[i] = bu(u_bin(op, u_(a.l[i]), u_(b.l[i])));
            break;
    }
    return r;
}

__device__ __forceinline__ V4 vm_un(uint32_t op, int prec, const V4& a)
{
    V4 r;
    switch (prec) {
        case CHAI_F32:
#pragma unroll
            for (int i = 0; i < 4; i++) r.l[i] = bf(fp_un<float>(op, f_(a.l[i])));
            break;
        case CHAI_F64:
#pragma unroll
            for (int i = 0; i < 4; i++) r.l[i] = bd(fp_un<double>(op, d_(a.l[i])));
            break;
        case CHAI_I32:
#pragma unroll
            for (int i = 0; i < 4; i++) r.l[i] = bi(i_un(op, i_(a.l[i])));
            break;
        default:
#pragma unroll
            for (int i = 0; i < 4; i++) r.l[i] = bu(u_un(op, u_(a.l[i])));
            break;
    }
    return r;
}

__device__ __forceinline__ V4 vm_ter(uint32_t op, int prec, const V4& a, const V4& b, const V4& c)
{
    V4 r;
    switch (prec) {
        case CHAI_F32:
#pragma unroll
            for (int i = 0; i < 4; i++) r.l[i] = bf(ter<float>(op, f_(a.l[i]), f_(b.l[i]), f_(c.l[i])));
            break;
        case CHAI_F64:
#pragma unroll
            for (int i = 0; i < 4; i++) r.l[i] = bd(ter<double>(op, d_(a.l[i]), d_(b.l[i]), d_(c.l[i])));
            break;
        case CHAI_I32:
#pragma unroll
            for (int i = 0; i < 4; i++) r.l[i] = bi(i_ter(op, i_(a.l[i]), i_(b.l[i]), i_(c.l[i])));
            break;
        default:
#pragma unroll
            for (int i = 0; i < 4; i++) r.l[i] = bu(u_ter(op, u_(a.l[i]), u_(b.l[i]), u_(c.l[i])));
            break;
    }
    return r;
}

struct VmLaunchArgs {
    const VmIns* prog;
    uint32_t nins, nacc;
    uint64_t W, H, S, vpr, vps;
    double* partial;
    VmVar vars[24];
    VmRngDims rng[8];
    VmChain chains[8];         // leaf instruction field p2 = chain index + 1 (0: none)
};


}  // namespace chai_hip
