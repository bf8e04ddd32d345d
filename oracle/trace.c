/*
 * trace.c -- oracle evaluation of a trace (include/chai_mi355x.h chai_trace)
 * with Chai's interpreter semantics.  TEST INFRASTRUCTURE.
 *
 * Restates:
 *   - one full pass per operation into a fresh array, as the interpreter
 *     does (interp/BaseInterp.cpp:152-183);
 *   - modulo broadcast of operands: element (x, y) of an operand is
 *     [(x % W) + (y % H) * W], slot s -> s % slots (interp/BaseInterp.cpp:50-55,
 *     BaseInterp.cpp:57-85);
 *   - result precision = max of the operand PrecType codes
 *     (interp/InterpFun2.cpp:21, misc/PrecType.hpp:18-21); mixed operands are
 *     converted as misc/SimpleFun2.cpp:40-110 does (uint32 <-> int32 by bit
 *     reinterpretation, integers to floating point by value);
 *   - scalar ops of misc/SimpleFun{1,2,3}.cpp, with the reference build's x86
 *     behaviour made explicit where C leaves it undefined: shift counts use
 *     their low five bits; integer division / modulo by zero (and INT_MIN / -1)
 *     yield 0 (the reference would trap);
 *   - cond: type and shape from the two branches (interp/InterpCond.cpp:13-80);
 *   - sum / mean / dot in a double accumulator (interp/InterpAccum.cpp,
 *     interp/InterpDotprod.cpp); mean divides by W*H;
 *   - matmul / matmulG with interpreter GEMM semantics (linalg.c);
 *   - rng_*_make with the canonical stream (rng.c);
 *   - gather1_floor / gather2_floor with the interpreter's subscript
 *     arithmetic (interp/InterpGather.hpp:20-52), x86 float -> int conversion.
 * Floating point transcendentals are glibc's (as the interpreter's); the GPU
 * uses the device math library, so those results are compared with a stated
 * ulp tolerance, not bit for bit.
 *
 * Pinned by: the integer scalar ops against the reference's own
 * misc/SimpleFun2.cpp / SimpleFun1.cpp compiled into oracle/_ref
 * (tests/test_oracle_ref.py); closed forms (MD5 of known messages as an
 * integer trace, sums of index ramps) in tests/test_oracle_trace.py; the
 * gathers against a periodic 5-point stencil (np.roll, the reference's
 * sample/filter.cpp workload), a binomial-lattice backward step (numpy
 * slicing, sample/finance/american_option.cpp) and the subscript rule's closed
 * form per range of the floored index, in tests/test_oracle_trace.py.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#include "chai_oracle.h"
#include "../include/chai_mi355x.h"

typedef struct {
    int prec;
    size_t W, H, S;
    void* d;          /* [S][H][W] */
} Arr;

static size_t esz(int p) { return p == 8 ? 8 : 4; }

static Arr arr_new(int prec, size_t W, size_t H, size_t S)
{
    Arr a;
    a.prec = prec; a.W = W; a.H = H; a.S = S;
    a.d = calloc(W * H * S, esz(prec));
    return a;
}

static double getd(const Arr* a, size_t i)
{
    switch (a->prec) {
        case 0: return (double)((uint32_t*)a->d)[i];
        case 1: return (double)((int32_t*)a->d)[i];
        case 4: return (double)((float*)a->d)[i];
        default: return ((double*)a->d)[i];
    }
}

static uint32_t getu(const Arr* a, size_t i)   /* 32-bit integer pattern */
{
    return ((uint32_t*)a->d)[i];
}

/* element of a broadcast operand at (x, y, s) */
static size_t bidx(const Arr* a, size_t x, size_t y, size_t s)
{
    return ((s % a->S) * a->H + (y % a->H)) * a->W + (x % a->W);
}

/* value of operand a converted to precision p, SimpleFun2 mixed-overload rules */
static void conv_to(const Arr* a, size_t i, int p, uint32_t* u, double* d)
{
    if (p == 0 || p == 1) {
        if (a->prec == 0 || a->prec == 1) *u = getu(a, i);
        else *u = (uint32_t)(int64_t)getd(a, i);
    } else {
        *d = getd(a, i);
        if (p == 4) *d = (double)(float)*d;
    }
}

static void put(Arr* r, size_t i, uint32_t u, double d)
{
    switch (r->prec) {
        case 0: ((uint32_t*)r->d)[i] = u; break;
        case 1: ((int32_t*)r->d)[i] = (int32_t)u; break;
        case 4: ((float*)r->d)[i] = (float)d; break;
        default: ((double*)r->d)[i] = d; break;
    }
}

static uint32_t shl32(uint32_t a, uint32_t n) { return a << (n & 31u); }
static uint32_t shr32u(uint32_t a, uint32_t n) { return a >> (n & 31u); }
static int32_t shr32s(int32_t a, uint32_t n) { return a >> (n & 31u); }

static int bin_int(uint32_t op, int is_signed, uint32_t ua, uint32_t ub, uint32_t* r)
{
    const int32_t a = (int32_t)ua, b = (int32_t)ub;
    if (!is_signed) {
        switch (op) {
            case 11: *r = ua + ub; return 0;
            case 12: *r = ub ? ua / ub : 0u; return 0;
            case 13: *r = ua * ub; return 0;
            case 14: *r = ua - ub; return 0;
            case 17: *r = ub ? ua % ub : 0u; return 0;
            case 18: *r = shl32(ua, ub); return 0;
            case 19: *r = shr32u(ua, ub); return 0;
            case 20: *r = ua & ub; return 0;
            case 21: *r = ua | ub; return 0;
            case 22: *r = ua ^ ub; return 0;
            case 30: *r = ua && ub; return 0;
            case 31: *r = ua || ub; return 0;
            case 32: *r = ua == ub; return 0;
            case 33: *r = ua != ub; return 0;
            case 34: *r = ua >= ub; return 0;
            case 35: *r = ua > ub; return 0;
            case 36: *r = ua <= ub; return 0;
            case 37: *r = ua < ub; return 0;
            case 302: *r = ua < ub ? ub : ua; return 0;
            case 303: *r = ub < ua ? ub : ua; return 0;
            case 401: *r = ua > ub ? ua - ub : ub - ua; return 0;
            case 402: { uint64_t c = (uint64_t)ua + ub; *r = c > 0xffffffffull ? 0xffffffffu : (uint32_t)c; return 0; }
            case 405: { uint64_t c = (uint64_t)ua + ub; *r = (c <= 0xffffffffull ? (uint32_t)c : 0xffffffffu) >> 1; return 0; }
            case 411: *r = (ua & 0x00ffffffu) * (ub & 0x00ffffffu); return 0;
            case 412: *r = (uint32_t)(((uint64_t)ua * ub) >> 32); return 0;
            case 413: { uint64_t c = (uint64_t)ua + ub + 1; *r = (uint32_t)((c <= 0xffffffffull ? c : 0xffffffffull) >> 1); return 0; }
            case 414: *r = shl32(ua, ub) | shr32u(ua, 32u - ub); return 0;
            case 415: *r = ua < ub ? 0u : ub - ua; return 0;   /* SimpleFun2.cpp:521-527 returns b - a */
        }
        return -1;
    }
    switch (op) {
        case 11: *r = ua + ub; return 0;
        case 12: *r = (b == 0 || (a == INT32_MIN && b == -1)) ? 0u : (uint32_t)(a / b); return 0;
        case 13: *r = ua * ub; return 0;
        case 14: *r = ua - ub; return 0;
        case 17: *r = (b == 0 || (a == INT32_MIN && b == -1)) ? 0u : (uint32_t)(a % b); return 0;
        case 18: *r = shl32(ua, ub); return 0;
        case 19: *r = (uint32_t)shr32s(a, ub); return 0;
        case 20: *r = ua & ub; return 0;
        case 21: *r = ua | ub; return 0;
        case 22: *r = ua ^ ub; return 0;
        case 30: *r = a && b; return 0;
        case 31: *r = a || b; return 0;
        case 32: *r = a == b; return 0;
        case 33: *r = a != b; return 0;
        case 34: *r = a >= b; return 0;
        case 35: *r = a > b; return 0;
        case 36: *r = a <= b; return 0;
        case 37: *r = a < b; return 0;
        case 302: *r = (uint32_t)(a < b ? b : a); return 0;
        case 303: *r = (uint32_t)(b < a ? b : a); return 0;
        case 401: *r = a > b ? ua - ub : ub - ua; return 0;
        /* upper saturation only, the low side wraps (SimpleFun2.cpp:187-191, 233-238) */
        case 402: { long long c = (long long)a + b; *r = (uint32_t)(c <= INT32_MAX ? (int32_t)(uint32_t)c : INT32_MAX); return 0; }
        case 405: { long long c = (long long)a + b; int32_t d = c <= INT32_MAX ? (int32_t)(uint32_t)c : INT32_MAX; *r = (uint32_t)(d >> 1); return 0; }
        case 411: *r = (uint32_t)(a & 0x00ffffff) * (uint32_t)(b & 0x00ffffff); return 0;
        case 412: *r = (uint32_t)(int32_t)(((long long)a * b) >> 32); return 0;
        case 413: { long long c = (long long)a + b + 1; int32_t d = c <= INT32_MAX ? (int32_t)(uint32_t)c : INT32_MAX; *r = (uint32_t)(d >> 1); return 0; }
        case 414: *r = shl32(ua, ub) | (uint32_t)shr32s(a, 32u - ub); return 0;
        case 415: { long long c = (long long)a - b; *r = (uint32_t)(int32_t)(c > INT32_MAX ? INT32_MAX : c < INT32_MIN ? INT32_MIN : c); return 0; }
    }
    return -1;
}

/* floating point binary op in precision p (4: float arithmetic, 8: double) */
/* pown in double: the reference's pow(double, int) call resolves, in the
 * C++98 library it is built with, to repeated squaring (x^|n| by
 * square-and-multiply from the low bit, reciprocal for n < 0) rather than to
 * pow(double, double); the float overload calls powf. */
static double powi_sq(double x, int m)
{
    unsigned n = m < 0 ? 0u - (unsigned)m : (unsigned)m;
    double y = (n & 1u) ? x : 1.0;
    while (n >>= 1) {
        x = x * x;
        if (n & 1u) y = y * x;
    }
    return m < 0 ? 1.0 / y : y;
}

/* signbit: in that build the float overload returns the raw sign bit
 * (0x80000000 as an int, i.e. -2^31 once converted back to float), the
 * double overload 1 -- pinned against oracle/_ref. */
static int bin_fp(uint32_t op, int p, double a, double b, double* r)
{
    if (p == 4) {
        const float fa = (float)a, fb = (float)b;
        float v;
        switch (op) {
            case 11: v = fa + fb; break;
            case 12: v = fa / fb; break;
            case 13: v = fa * fb; break;
            case 14: v = fa - fb; break;
            case 32: v = fa == fb; break;
            case 33: v = fa != fb; break;
            case 34: v = fa >= fb; break;
            case 35: v = fa > fb; break;
            case 36: v = fa <= fb; break;
            case 37: v = fa < fb; break;
            case 302: v = fa < fb ? fb : fa; break;
            case 303: v = fa < fb ? fa : fb; break;     /* SimpleFun2.cpp:349-356 */
            case 227: v = fmaxf(fa, fb); break;
            case 228: v = fminf(fa, fb); break;
            case 248: v = powf(fa, fb); break;
            case 207: v = atan2f(fa, fb); break;
            case 229: v = fmodf(fa, fb); break;
            case 232: v = hypotf(fa, fb); break;
            case 213: v = copysignf(fa, fb); break;
            case 224: v = fdimf(fa, fb); break;
            case 210: v = (float)(atan2f(fa, fb) / M_PI); break;
            case 234: v = ldexpf(fa, (int)fb); break;
            case 243: v = fabsf(fa) > fabsf(fb) ? fa : fabsf(fa) < fabsf(fb) ? fb : fmaxf(fa, fb); break;
            case 244: v = fabsf(fa) < fabsf(fb) ? fa : fabsf(fa) > fabsf(fb) ? fb : fminf(fa, fb); break;
            case 247: v = nextafterf(fa, fb); break;
            case 249: v = powf(fa, (float)(int)fb); break;
            case 250: v = powf(fa, fb); break;
            case 251: v = remainderf(fa, fb); break;
            case 254: v = powf(fa, 1.f / fb); break;
            case 306: v = fb < fa ? 0.0f : 1.0f; break;
            case 500: v = fa == fb; break;
            case 501: v = fa != fb; break;
            case 502: v = fa > fb; break;
            case 503: v = fa >= fb; break;
            case 504: v = fa < fb; break;
            case 505: v = fa <= fb; break;
            case 506: v = (fa < fb) || (fa > fb); break;
            case 511: v = fa == fa && fb == fb; break;
            case 512: v = isnan(fa) || isnan(fb); break;
            default: return -1;
        }
        *r = v;
        return 0;
    }
    switch (op) {
        case 11: *r = a + b; return 0;
        case 12: *r = a / b; return 0;
        case 13: *r = a * b; return 0;
        case 14: *r = a - b; return 0;
        case 32: *r = a == b; return 0;
        case 33: *r = a != b; return 0;
        case 34: *r = a >= b; return 0;
        case 35: *r = a > b; return 0;
        case 36: *r = a <= b; return 0;
        case 37: *r = a < b; return 0;
        case 302: *r = a < b ? b : a; return 0;
        case 303: *r = a < b ? a : b; return 0;
        case 227: *r = fmax(a, b); return 0;
        case 228: *r = fmin(a, b); return 0;
        case 248: *r = pow(a, b); return 0;
        case 207: *r = atan2(a, b); return 0;
        case 229: *r = fmod(a, b); return 0;
        case 232: *r = hypot(a, b); return 0;
        case 213: *r = copysign(a, b); return 0;
        case 224: *r = fdim(a, b); return 0;
        case 210: *r = atan2(a, b) / M_PI; return 0;
        case 234: *r = ldexp(a, (int)b); return 0;
        case 243: *r = fabs(a) > fabs(b) ? a : fabs(a) < fabs(b) ? b : fmax(a, b); return 0;
        case 244: *r = fabs(a) < fabs(b) ? a : fabs(a) > fabs(b) ? b : fmin(a, b); return 0;
        case 247: *r = nextafter(a, b); return 0;
        case 249: *r = powi_sq(a, (int)b); return 0;
        case 250: *r = pow(a, b); return 0;
        case 251: *r = remainder(a, b); return 0;
        case 254: *r = pow(a, 1.0 / b); return 0;
        case 306: *r = b < a ? 0.0 : 1.0; return 0;
        case 500: *r = a == b; return 0;
        case 501: *r = a != b; return 0;
        case 502: *r = a > b; return 0;
        case 503: *r = a >= b; return 0;
        case 504: *r = a < b; return 0;
        case 505: *r = a <= b; return 0;
        case 506: *r = (a < b) || (a > b); return 0;
        case 511: *r = a == a && b == b; return 0;
        case 512: *r = isnan(a) || isnan(b); return 0;
    }
    return -1;
}

static int un_fp(uint32_t op, int p, double a, double* r)
{
    if (p == 4) {
        const float f = (float)a;
        float v;
        switch (op) {
            case 10: v = -f; break;
            case 223: v = fabsf(f); break;
            case 261: v = sqrtf(f); break;
            case 256: v = 1.0f / sqrtf(f); break;
            case 219: v = expf(f); break;
            case 220: v = exp2f(f); break;
            case 221: v = exp10f(f); break;
            case 222: v = expm1f(f); break;
            case 237: v = logf(f); break;
            case 238: v = log2f(f); break;
            case 239: v = log10f(f); break;
            case 240: v = log1pf(f); break;
            case 257: v = sinf(f); break;
            case 214: v = cosf(f); break;
            case 262: v = tanf(f); break;
            case 203: v = asinf(f); break;
            case 200: v = acosf(f); break;
            case 206: v = atanf(f); break;
            case 259: v = sinhf(f); break;
            case 215: v = coshf(f); break;
            case 263: v = tanhf(f); break;
            case 225: v = floorf(f); break;
            case 212: v = ceilf(f); break;
            case 253: v = rintf(f); break;
            case 255: v = roundf(f); break;
            case 266: v = truncf(f); break;
            case 211: v = cbrtf(f); break;
            case 218: v = erff(f); break;
            case 217: v = erfcf(f); break;
            /* SimpleFun1.cpp: the pi variants and degrees / radians go through double */
            case 201: v = acoshf(f); break;
            case 202: v = (float)(acosf(f) / M_PI); break;
            case 204: v = asinhf(f); break;
            case 205: v = (float)(asinf(f) / M_PI); break;
            case 208: v = atanhf(f); break;
            case 209: v = (float)(atanf(f) / M_PI); break;
            case 216: v = cosf((float)(M_PI * f)); break;
            case 233: v = (float)ilogbf(f); break;
            case 235: v = lgammaf(f); break;
            case 241: v = logbf(f); break;
            case 246: v = (float)nan(""); break;
            case 260: v = sinf((float)(M_PI * f)); break;
            case 264: v = tanf((float)(M_PI * f)); break;
            case 265: v = tgammaf(f); break;
            case 301: v = (float)(M_1_PI * 180 * f); break;
            case 305: v = (float)(f * M_PI / 180); break;
            case 308: v = f > 0.0 ? 1.0f : f < 0.0 ? -1.0f : f; break;
            case 507: v = isfinite(f) ? 1.0f : 0.0f; break;
            case 508: v = isinf(f) ? 1.0f : 0.0f; break;
            case 509: v = isnan(f) ? 1.0f : 0.0f; break;
            case 510: v = isnormal(f) ? 1.0f : 0.0f; break;
            case 513: v = signbit(f) ? -2147483648.0f : 0.0f; break;   /* see below */
            default: return -1;
        }
        *r = v;
        return 0;
    }
    switch (op) {
        case 10: *r = -a; return 0;
        case 223: *r = fabs(a); return 0;
        case 261: *r = sqrt(a); return 0;
        case 256: *r = 1.0 / sqrt(a); return 0;
        case 219: *r = exp(a); return 0;
        case 220: *r = exp2(a); return 0;
        case 221: *r = exp10(a); return 0;
        case 222: *r = expm1(a); return 0;
        case 237: *r = log(a); return 0;
        case 238: *r = log2(a); return 0;
        case 239: *r = log10(a); return 0;
        case 240: *r = log1p(a); return 0;
        case 257: *r = sin(a); return 0;
        case 214: *r = cos(a); return 0;
        case 262: *r = tan(a); return 0;
        case 203: *r = asin(a); return 0;
        case 200: *r = acos(a); return 0;
        case 206: *r = atan(a); return 0;
        case 259: *r = sinh(a); return 0;
        case 215: *r = cosh(a); return 0;
        case 263: *r = tanh(a); return 0;
        case 225: *r = floor(a); return 0;
        case 212: *r = ceil(a); return 0;
        case 253: *r = rint(a); return 0;
        case 255: *r = round(a); return 0;
        case 266: *r = trunc(a); return 0;
        case 211: *r = cbrt(a); return 0;
        case 218: *r = erf(a); return 0;
        case 217: *r = erfc(a); return 0;
        case 201: *r = acosh(a); return 0;
        case 202: *r = acos(a) / M_PI; return 0;
        case 204: *r = asinh(a); return 0;
        case 205: *r = asin(a) / M_PI; return 0;
        case 208: *r = atanh(a); return 0;
        case 209: *r = atan(a) / M_PI; return 0;
        case 216: *r = cos(M_PI * a); return 0;
        case 233: *r = (double)ilogb(a); return 0;
        case 235: *r = lgamma(a); return 0;
        case 241: *r = logb(a); return 0;
        case 246: *r = nan(""); return 0;
        case 260: *r = sin(M_PI * a); return 0;
        case 264: *r = tan(M_PI * a); return 0;
        case 265: *r = tgamma(a); return 0;
        case 301: *r = M_1_PI * 180 * a; return 0;
        case 305: *r = a * M_PI / 180; return 0;
        case 308: *r = a > 0.0 ? 1.0 : a < 0.0 ? -1.0 : a; return 0;
        case 507: *r = isfinite(a) ? 1.0 : 0.0; return 0;
        case 508: *r = isinf(a) ? 1.0 : 0.0; return 0;
        case 509: *r = isnan(a) ? 1.0 : 0.0; return 0;
        case 510: *r = isnormal(a) ? 1.0 : 0.0; return 0;
        case 513: *r = signbit(a) ? 1.0 : 0.0; return 0;
    }
    return -1;
}

static int un_int(uint32_t op, int is_signed, uint32_t a, uint32_t* r)
{
    switch (op) {
        case 10: *r = 0u - a; return 0;
        case 15: *r = ~a; return 0;
        case 16: *r = !a; return 0;
        case 400: *r = (is_signed && (int32_t)a < 0) ? 0u - a : a; return 0;
        case 404: *r = a ? (uint32_t)__builtin_clz(a) : 32u; return 0;
        case 225: *r = a; return 0;
    }
    return -1;
}

#define MAXSTK 64

static size_t maxz(size_t a, size_t b) { return a > b ? a : b; }

/* gather subscript, step 1 (interp/InterpGather.hpp:24, misc/SimpleFun1.cpp:196-216):
 * floor the element (identity for the integer types) and convert to int the
 * way the reference's x86 build does (cvttss2si / cvttsd2si: NaN and values
 * outside int give INT_MIN; uint32 reinterprets its bits). */
static int gather_floor(const Arr* a, size_t i)
{
    double v;
    switch (a->prec) {
        case 0: case 1: return (int)(int32_t)getu(a, i);
        case 4: v = (double)floorf(((const float*)a->d)[i]); break;
        default: v = floor(((const double*)a->d)[i]); break;
    }
    return (v != v || v >= 2147483648.0 || v < -2147483648.0) ? INT32_MIN : (int)v;
}

/* gather subscript, step 2 (interp/InterpGather.hpp:25-26): the interpreter's
 * "(xf < 0) ? xf + W : xf" mixes int and size_t, so it is computed in size_t
 * and stored back to an int; "xfM % W" then converts that int to size_t
 * (sign-extending) before an unsigned 64-bit modulo. */
static size_t gather_wrap(int xf, size_t W)
{
    const size_t m = xf < 0 ? (size_t)(ptrdiff_t)xf + W : (size_t)(ptrdiff_t)xf;
    const int xfM = (int)(uint32_t)m;
    return (size_t)(ptrdiff_t)xfM % W;
}

int chai_or_trace_eval(const chai_trace* t, void* const* bufs)
{
    Arr stk[MAXSTK];
    uint32_t root[MAXSTK];     /* node index of each stack entry's root (matmul transpose flags) */
    int sp = 0;
    int rc = 0;
    for (uint32_t si = 0; si < t->num_stmts && !rc; si++) {
        const chai_trace_stmt* st = &t->stmts[si];
        sp = 0;
        for (uint32_t ni = 0; ni < st->count && !rc; ni++) {
            const chai_trace_node* n = &t->nodes[st->first + ni];
            const uint32_t c = n->code;
            Arr r;
            if (sp >= MAXSTK - 1) return -1;
            if (c == CHAI_TRACE_VAR) {
                const chai_trace_var* v = &t->vars[n->arg];
                r = arr_new((int)v->prec, v->W, v->H, v->slots);
                memcpy(r.d, bufs[n->arg], esz(r.prec) * v->W * v->H * v->slots);
            } else if (c >= 5 && c <= 8) {
                static const int sp4[] = {0, 1, 4, 8};
                r = arr_new(sp4[c - 5], 1, 1, 1);
                if (r.prec == 0) ((uint32_t*)r.d)[0] = (uint32_t)(int64_t)n->s0;
                else if (r.prec == 1) ((int32_t*)r.d)[0] = (int32_t)(int64_t)n->s0;
                else put(&r, 0, 0, n->s0);
            } else if (c >= 110 && c <= 115) {
                static const int rp[] = {4, 8, 0, 1, 4, 8};
                const size_t H = n->H ? n->H : 1, S = n->slots ? n->slots : 1;
                r = arr_new(rp[c - 110], n->W, H, S);
                const int normal = c <= 111;
                const int lim = !normal && (c == 114 || c == 115) && (n->s0 != 0.0 || n->s1 != 0.0);
                if (chai_or_rng_fill((int)n->arg, r.prec, (int)n->vec_len, n->seed, n->W, H, S, normal,
                                     lim, n->s0, n->s1, r.d))
                    rc = -1;
            } else if (c >= 68 && c <= 82) {
                /* ones / zeros (runtime/Translator.cpp:116-131, jitast/AstLitdata.cpp):
                 * the translator's typed values; 73 is declared for both ones2_i32
                 * and zeros2_f64 (bytecode/ByteCodes.hpp:95,105) and the dispatch
                 * table keeps the later registration, zeros2_f64 */
                static const int lp[] = {0, 1, 4, 8};
                const int p = c == 73 ? 8 : lp[(c - 68) % 4];
                const double v = (c == 73 || c >= 76) ? 0.0 : 1.0;
                const size_t H = n->H ? n->H : 1;
                r = arr_new(p, n->W, H, 1);
                for (size_t i = 0; i < n->W * H; i++) put(&r, i, (uint32_t)v, v);
            } else if (c >= 42 && c <= 49) {
                static const int ip[] = {0, 1, 4, 8};
                const size_t H = n->H ? n->H : 1, S = n->slots ? n->slots : 1;
                r = arr_new(ip[(c - 42) % 4], n->W, H, S);
                for (size_t s = 0; s < S; s++)
                    for (size_t y = 0; y < H; y++)
                        for (size_t x = 0; x < n->W; x++) {
                            const double v = n->arg ? (double)y : (double)x;
                            put(&r, (s * H + y) * n->W + x, (uint32_t)(int64_t)v, v);
                        }
            } else if (c >= 1 && c <= 4) {
                static const int cp[] = {0, 1, 4, 8};
                Arr a = stk[--sp];
                r = arr_new(cp[c - 1], a.W, a.H, a.S);
                for (size_t i = 0; i < a.W * a.H * a.S; i++) {
                    const double d = getd(&a, i);   /* conversion through double */
                    put(&r, i, (uint32_t)(int64_t)d, d);
                }
                free(a.d);
            } else if (c == 122) {                         /* transpose (interp/InterpTranspose.cpp:13-60) */
                Arr a = stk[--sp];
                r = arr_new(a.prec, a.H, a.W, a.S);
                const size_t e = esz(a.prec);
                for (size_t s = 0; s < a.S; s++)
                    for (size_t y = 0; y < a.H; y++)
                        for (size_t x = 0; x < a.W; x++)
                            memcpy((char*)r.d + ((s * a.W + x) * a.H + y) * e,
                                   (const char*)a.d + ((s * a.H + y) * a.W + x) * e, e);
                free(a.d);
            } else if (c == 131 || c == 132) {
                Arr a = stk[--sp];
                r = arr_new(a.prec, 1, 1, a.S);
                if (chai_or_reduce((int)c, a.prec, a.d, a.W * a.H, a.S, r.d)) rc = -1;
                free(a.d);
            } else if (c == 130) {
                Arr b = stk[--sp], a = stk[--sp];
                const int p = a.prec > b.prec ? a.prec : b.prec;
                const size_t W = maxz(a.W, b.W), H = maxz(a.H, b.H), S = maxz(a.S, b.S);
                r = arr_new(p, 1, 1, S);
                for (size_t s = 0; s < S; s++) {
                    double acc = 0.0;
                    for (size_t y = 0; y < H; y++)
                        for (size_t x = 0; x < W; x++) {
                            uint32_t ua = 0, ub = 0, ur = 0;
                            double da = 0, db = 0, dr = 0;
                            conv_to(&a, bidx(&a, x, y, s), p, &ua, &da);
                            conv_to(&b, bidx(&b, x, y, s), p, &ub, &db);
                            if (p <= 1) { bin_int(13, p == 1, ua, ub, &ur); acc += p == 1 ? (double)(int32_t)ur : (double)ur; }
                            else { bin_fp(13, p, da, db, &dr); acc += dr; }
                        }
                    put(&r, s, (uint32_t)(int64_t)acc, acc);
                }
                free(a.d); free(b.d);
            } else if (c == 120 || c == 121) {
                const int general = c == 121;
                Arr C0, B, A;
                if (general) C0 = stk[--sp];
                B = stk[--sp];
                A = stk[--sp];
                const chai_trace_node* na = &t->nodes[root[sp]];
                const chai_trace_node* nb = &t->nodes[root[sp + 1]];
                const int ta = na->flags & 1, tb = nb->flags & 1;
                const int p = A.prec > B.prec ? A.prec : B.prec;
                /* slots (interp/InterpMatmulBase.cpp:113-118, BaseInterp.cpp:79-98):
                 * max(A.S, B.S) products, C's count for matmulG; instance s
                 * reads A's instance s mod A.S and B's s mod B.S */
                const size_t S = general ? C0.S : maxz(A.S, B.S);
                const size_t ea = A.W * A.H * esz(A.prec), eb = B.W * B.H * esz(B.prec);
                size_t M, N, K = 0;
                int kind;                                   /* 0 mat*vec, 1 vec*mat, 2 mat*mat */
                if (B.H == 1 && !tb && A.H > 1) { kind = 0; M = ta ? A.W : A.H; N = ta ? A.H : A.W; }
                else if (A.H == 1 && !ta && B.H > 1) { kind = 1; M = tb ? B.H : B.W; N = tb ? B.W : B.H; }
                else { kind = 2; M = ta ? A.W : A.H; K = ta ? A.H : A.W; N = tb ? B.H : B.W; }
                if (general) {
                    r = C0;
                } else {
                    r = kind == 2 ? arr_new(p, N, M, S) : arr_new(p, M, 1, S);
                    C0.d = NULL;
                }
                const size_t ec = r.W * r.H * esz(r.prec);
                for (size_t q = 0; q < S && rc == 0; q++) {
                    const char* pa = (const char*)A.d + (q % A.S) * ea;
                    const char* pb = (const char*)B.d + (q % B.S) * eb;
                    char* pcq = (char*)r.d + q * ec;
                    if (kind == 0) {
                        if (chai_or_matvec_interp(A.prec, B.prec, r.prec, ta, M, N, pa, A.W, pb, pcq,
                                                  general, n->s0, n->s1)) rc = -1;
                    } else if (kind == 1) {
                        if (chai_or_matvec_interp(B.prec, A.prec, r.prec, !tb, M, N, pb, B.W, pa, pcq,
                                                  general, n->s0, n->s1)) rc = -1;
                    } else {
                        if (chai_or_matmul_interp(A.prec, B.prec, r.prec, ta, tb, M, N, K, pa, A.W, pb,
                                                  B.W, pcq, N, general, n->s0, n->s1)) rc = -1;
                    }
                }
                free(A.d); free(B.d);
            } else if (c == 40 || c == 41) {
                /* gather1_floor / gather2_floor (interp/InterpGather.cpp:20-60,
                 * InterpGather.hpp:20-52): shape and type of the data array;
                 * subscripts read by modulo broadcast at (x, y, slot) */
                Arr iy = {0, 0, 0, 0, NULL}, ix, a;
                if (c == 41) iy = stk[--sp];
                ix = stk[--sp];
                a = stk[--sp];
                r = arr_new(a.prec, a.W, a.H, a.S);
                const size_t e = esz(a.prec);
                for (size_t s = 0; s < a.S; s++)
                    for (size_t y = 0; y < a.H; y++)
                        for (size_t x = 0; x < a.W; x++) {
                            const size_t xr = gather_wrap(gather_floor(&ix, bidx(&ix, x, y, s)), a.W);
                            const size_t yr = c == 41 ? gather_wrap(gather_floor(&iy, bidx(&iy, x, y, s)), a.H) : y;
                            memcpy((char*)r.d + ((s * a.H + y) * a.W + x) * e,
                                   (const char*)a.d + ((s * a.H + yr) * a.W + xr) * e, e);
                        }
                free(a.d); free(ix.d);
                if (c == 41) free(iy.d);
            } else if (c == 23) {                          /* cond */
                Arr b2 = stk[--sp], b1 = stk[--sp], cc = stk[--sp];
                const int p = b1.prec > b2.prec ? b1.prec : b2.prec;
                const size_t W = maxz(b1.W, b2.W), H = maxz(b1.H, b2.H), S = maxz(b1.S, b2.S);
                r = arr_new(p, W, H, S);
                for (size_t s = 0; s < S; s++)
                    for (size_t y = 0; y < H; y++)
                        for (size_t x = 0; x < W; x++) {
                            const Arr* src = getd(&cc, bidx(&cc, x, y, s)) != 0.0 ? &b1 : &b2;
                            const double d = getd(src, bidx(src, x, y, s));
                            uint32_t u = (src->prec <= 1 && p <= 1) ? getu(src, bidx(src, x, y, s))
                                                                    : (uint32_t)(int64_t)d;
                            put(&r, (s * H + y) * W + x, u, d);
                        }
                free(cc.d); free(b1.d); free(b2.d);
            } else {
                /* unary / binary / ternary elementwise */
                int arity;
                switch (c) {
                    case 10: case 15: case 16: case 400: case 404: case 225: case 223: case 261: case 256:
                    case 219: case 220: case 221: case 222: case 237: case 238: case 239: case 240: case 257:
                    case 214: case 262: case 203: case 200: case 206: case 259: case 215: case 263: case 212:
                    case 253: case 255: case 266: case 211: case 218: case 217:
                    case 201: case 202: case 204: case 205: case 208: case 209: case 216: case 233:
                    case 235: case 241: case 246: case 260: case 264: case 265: case 301: case 305:
                    case 308: case 507: case 508: case 509: case 510: case 513:
                        arity = 1; break;
                    case 300: case 226: case 242: case 304: case 307: case 406: case 407: case 408:
                        arity = 3; break;
                    default:
                        arity = 2; break;
                }
                Arr k[3];
                for (int j = arity - 1; j >= 0; j--) k[j] = stk[--sp];
                int p = k[0].prec;
                size_t W = k[0].W, H = k[0].H, S = k[0].S;
                for (int j = 1; j < arity; j++) {
                    if (k[j].prec > p) p = k[j].prec;
                    W = maxz(W, k[j].W); H = maxz(H, k[j].H); S = maxz(S, k[j].S);
                }
                r = arr_new(p, W, H, S);
                for (size_t s = 0; s < S && !rc; s++)
                    for (size_t y = 0; y < H && !rc; y++)
                        for (size_t x = 0; x < W; x++) {
                            uint32_t u[3] = {0, 0, 0}, ru = 0;
                            double d[3] = {0, 0, 0}, rd = 0;
                            for (int j = 0; j < arity; j++) conv_to(&k[j], bidx(&k[j], x, y, s), p, &u[j], &d[j]);
                            int e;
                            if (arity == 1) {
                                e = p <= 1 ? un_int(c, p == 1, u[0], &ru) : un_fp(c, p, d[0], &rd);
                            } else if (arity == 2) {
                                e = p <= 1 ? bin_int(c, p == 1, u[0], u[1], &ru) : bin_fp(c, p, d[0], d[1], &rd);
                            } else {
                                e = 0;
                                if (c == 300) {
                                    if (p == 0) { uint32_t m = u[0] < u[1] ? u[1] : u[0]; ru = m < u[2] ? m : u[2]; }
                                    else if (p == 1) { int32_t a0 = (int32_t)u[0], a1 = (int32_t)u[1], a2 = (int32_t)u[2];
                                                       int32_t m = a0 < a1 ? a1 : a0; ru = (uint32_t)(m < a2 ? m : a2); }
                                    else if (p == 4) { float m = (float)d[0] < (float)d[1] ? (float)d[1] : (float)d[0];
                                                       rd = m < (float)d[2] ? m : (float)d[2]; }
                                    else { double m = d[0] < d[1] ? d[1] : d[0]; rd = m < d[2] ? m : d[2]; }
                                } else if (c == 226 && p >= 4) {
                                    rd = p == 4 ? (double)fmaf((float)d[0], (float)d[1], (float)d[2]) : fma(d[0], d[1], d[2]);
                                } else if (c == 242 && p >= 4) {
                                    if (p == 4) { float m = (float)d[0] * (float)d[1]; rd = m + (float)d[2]; }
                                    else { double m = d[0] * d[1]; rd = m + d[2]; }
                                } else if (c == 304 && p >= 4) {        /* mix: a + (b - a) * c */
                                    if (p == 4) { float t1 = (float)d[1] - (float)d[0]; float t2 = t1 * (float)d[2];
                                                  rd = (float)d[0] + t2; }
                                    else rd = d[0] + (d[1] - d[0]) * d[2];
                                } else if (c == 307 && p >= 4) {        /* smoothstep(edge0, edge1, x) */
                                    if (p == 4) {
                                        const float a0 = (float)d[0], b0 = (float)d[1], x0 = (float)d[2];
                                        const float q = (x0 - a0) / (b0 - a0);
                                        const float q0 = q < 0.0 ? 0.0f : q, tt = q0 < 1.0 ? q0 : 1.0f;
                                        const float t2 = tt * tt, u3 = 3 - 2 * tt;
                                        rd = t2 * u3;
                                    } else {
                                        const double q = (d[2] - d[0]) / (d[1] - d[0]);
                                        const double q0 = q < 0.0 ? 0.0 : q, tt = q0 < 1.0 ? q0 : 1.0;
                                        rd = tt * tt * (3 - 2 * tt);
                                    }
                                } else if ((c == 406 || c == 407) && p <= 1) {   /* mad24 / mad_hi */
                                    uint32_t m;
                                    bin_int(c == 406 ? 411 : 412, p == 1, u[0], u[1], &m);
                                    ru = m + u[2];
                                } else if (c == 408 && p <= 1) {                 /* mad_sat */
                                    if (p == 0) {
                                        const uint64_t dd = (uint64_t)u[0] * u[1] + u[2];
                                        ru = dd > 0xffffffffull ? 0xffffffffu : (uint32_t)dd;
                                    } else {
                                        const long long dd = (long long)(int32_t)u[0] * (int32_t)u[1] + (int32_t)u[2];
                                        ru = (uint32_t)(int32_t)(dd < INT32_MIN ? INT32_MIN : dd > INT32_MAX ? INT32_MAX : dd);
                                    }
                                } else {
                                    e = -1;
                                }
                            }
                            if (e) { rc = -1; break; }
                            put(&r, (s * H + y) * W + x, ru, rd);
                        }
                for (int j = 0; j < arity; j++) free(k[j].d);
            }
            root[sp] = st->first + ni;
            stk[sp++] = r;
        }
        if (rc) break;
        if (sp != 1) { rc = -1; break; }
        /* assignment converts to the variable's precision */
        const chai_trace_var* lv = &t->vars[st->lhs];
        Arr v = stk[0];
        const size_t n = lv->W * lv->H * lv->slots;
        if (v.W * v.H * v.S != n) { free(v.d); sp = 0; rc = -1; break; }
        if ((int)lv->prec == v.prec) {
            memcpy(bufs[st->lhs], v.d, esz(v.prec) * n);
        } else {
            Arr o; o.prec = (int)lv->prec; o.W = lv->W; o.H = lv->H; o.S = lv->slots; o.d = bufs[st->lhs];
            for (size_t i = 0; i < n; i++) { const double d = getd(&v, i); put(&o, i, (uint32_t)(int64_t)d, d); }
        }
        free(v.d);
        sp = 0;
    }
    for (int i = 0; i < sp; i++) free(stk[i].d);
    return rc;
}
