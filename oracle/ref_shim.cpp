// ref_shim.cpp -- C entry points onto the reference's own misc/ sources
// (misc/TEA.cpp, misc/SimpleFun{1,2,3}.cpp, misc/PrecType.cpp), compiled
// where they lie by `make -C oracle ref` into oracle/_ref/libchai_ref.so.
// TEST INFRASTRUCTURE: used only by tests/ to pin the oracle's restatement of
// the trace hash and of the scalar op semantics.  This file contains no
// reference code; it only forwards calls.
#include <cstring>
#include <stdint.h>
#include "TEA.hpp"
#include "PrecType.hpp"
#include "SimpleFun1.hpp"
#include "SimpleFun2.hpp"
#include "SimpleFun3.hpp"

using namespace chai_internal;

extern "C" {

uint64_t chai_ref_tea_hash(const uint64_t* data, size_t n)
{
    return TEA::hash(data, n);
}

size_t chai_ref_pad_width(size_t prec, size_t W)
{
    return PrecType::padWidth(prec, W);
}

static const SimpleFun2* fun2(int code)
{
    switch (code) {
        case 11: return &_ADD_::s();
        case 12: return &_DIVIDE_::s();
        case 13: return &_MULTIPLY_::s();
        case 14: return &_SUBTRACT_::s();
        case 17: return &_MODULO_::s();
        case 18: return &_SHIFTLEFT_::s();
        case 19: return &_SHIFTRIGHT_::s();
        case 20: return &_BITWISEAND_::s();
        case 21: return &_BITWISEOR_::s();
        case 22: return &_BITWISEXOR_::s();
        case 30: return &_LOGICALAND_::s();
        case 31: return &_LOGICALOR_::s();
        case 32: return &_EQUAL_::s();
        case 33: return &_NOTEQUAL_::s();
        case 34: return &_GREATEREQUAL_::s();
        case 35: return &_GREATERTHAN_::s();
        case 36: return &_LESSEQUAL_::s();
        case 37: return &_LESSTHAN_::s();
        case 302: return &_MAX_::s();
        case 303: return &_MIN_::s();
        case 401: return &_ABS_DIFF_::s();
        case 402: return &_ADD_SAT_::s();
        case 405: return &_HADD_::s();
        case 411: return &_MUL24_::s();
        case 412: return &_MUL_HI_::s();
        case 413: return &_RHADD_::s();
        case 414: return &_ROTATE_::s();
        case 415: return &_SUB_SAT_::s();
        case 207: return &_ATAN2_::s();
        case 210: return &_ATAN2PI_::s();
        case 213: return &_COPYSIGN_::s();
        case 224: return &_FDIM_::s();
        case 227: return &_FMAX_::s();
        case 228: return &_FMIN_::s();
        case 229: return &_FMOD_::s();
        case 232: return &_HYPOT_::s();
        case 234: return &_LDEXP_::s();
        case 243: return &_MAXMAG_::s();
        case 244: return &_MINMAG_::s();
        case 247: return &_NEXTAFTER_::s();
        case 248: return &_POW_::s();
        case 249: return &_POWN_::s();
        case 250: return &_POWR_::s();
        case 251: return &_REMAINDER_::s();
        case 254: return &_ROOTN_::s();
        case 306: return &_STEP_::s();
        case 500: return &_ISEQUAL_::s();
        case 501: return &_ISNOTEQUAL_::s();
        case 502: return &_ISGREATER_::s();
        case 503: return &_ISGREATEREQUAL_::s();
        case 504: return &_ISLESS_::s();
        case 505: return &_ISLESSEQUAL_::s();
        case 506: return &_ISLESSGREATER_::s();
        case 511: return &_ISORDERED_::s();
        case 512: return &_ISUNORDERED_::s();
        default: return 0;
    }
}

static const SimpleFun1* fun1(int code)
{
    switch (code) {
        case 10: return &_UNARYMINUS_::s();
        case 15: return &_COMPLEMENT_::s();
        case 16: return &_NOT_::s();
        case 400: return &_ABS_::s();
        case 404: return &_CLZ_::s();
        case 225: return &_FLOOR_::s();
        case 200: return &_ACOS_::s();
        case 201: return &_ACOSH_::s();
        case 202: return &_ACOSPI_::s();
        case 203: return &_ASIN_::s();
        case 204: return &_ASINH_::s();
        case 205: return &_ASINPI_::s();
        case 206: return &_ATAN_::s();
        case 208: return &_ATANH_::s();
        case 209: return &_ATANPI_::s();
        case 211: return &_CBRT_::s();
        case 212: return &_CEIL_::s();
        case 214: return &_COS_::s();
        case 215: return &_COSH_::s();
        case 216: return &_COSPI_::s();
        case 217: return &_ERFC_::s();
        case 218: return &_ERF_::s();
        case 219: return &_EXP_::s();
        case 220: return &_EXP2_::s();
        case 221: return &_EXP10_::s();
        case 222: return &_EXPM1_::s();
        case 223: return &_FABS_::s();
        case 233: return &_ILOGB_::s();
        case 235: return &_LGAMMA_::s();
        case 237: return &_LOG_::s();
        case 238: return &_LOG2_::s();
        case 239: return &_LOG10_::s();
        case 240: return &_LOG1P_::s();
        case 241: return &_LOGB_::s();
        case 246: return &_NAN_::s();
        case 253: return &_RINT_::s();
        case 255: return &_ROUND_::s();
        case 256: return &_RSQRT_::s();
        case 257: return &_SIN_::s();
        case 259: return &_SINH_::s();
        case 260: return &_SINPI_::s();
        case 261: return &_SQRT_::s();
        case 262: return &_TAN_::s();
        case 263: return &_TANH_::s();
        case 264: return &_TANPI_::s();
        case 265: return &_TGAMMA_::s();
        case 266: return &_TRUNC_::s();
        case 301: return &_DEGREES_::s();
        case 305: return &_RADIANS_::s();
        case 308: return &_SIGN_::s();
        case 507: return &_ISFINITE_::s();
        case 508: return &_ISINF_::s();
        case 509: return &_ISNAN_::s();
        case 510: return &_ISNORMAL_::s();
        case 513: return &_SIGNBIT_::s();
        default: return 0;
    }
}

static const SimpleFun3* fun3(int code)
{
    switch (code) {
        case 226: return &_FMA_::s();
        case 242: return &_MAD_::s();
        case 300: return &_CLAMP_::s();
        case 304: return &_MIX_::s();
        case 307: return &_SMOOTHSTEP_::s();
        case 406: return &_MAD24_::s();
        case 407: return &_MAD_HI_::s();
        case 408: return &_MAD_SAT_::s();
        default: return 0;
    }
}

// integer binary op: prec 0 = uint32, 1 = int32.  Returns 0 on success.
int chai_ref_fun2_int(int code, int prec, const uint32_t* a, const uint32_t* b,
                      uint32_t* out, size_t n)
{
    const SimpleFun2* f = fun2(code);
    if (!f) return -1;
    for (size_t i = 0; i < n; i++) {
        if (prec == 0) {
            out[i] = (*f)(a[i], b[i]);
        } else {
            int32_t x, y, r;
            std::memcpy(&x, &a[i], 4);
            std::memcpy(&y, &b[i], 4);
            r = (*f)(x, y);
            std::memcpy(&out[i], &r, 4);
        }
    }
    return 0;
}

int chai_ref_fun2_f32(int code, const float* a, const float* b, float* out, size_t n)
{
    const SimpleFun2* f = fun2(code);
    if (!f) return -1;
    for (size_t i = 0; i < n; i++) out[i] = (*f)(a[i], b[i]);
    return 0;
}

int chai_ref_fun1_int(int code, int prec, const uint32_t* a, uint32_t* out, size_t n)
{
    const SimpleFun1* f = fun1(code);
    if (!f) return -1;
    for (size_t i = 0; i < n; i++) {
        if (prec == 0) {
            out[i] = (*f)(a[i]);
        } else {
            int32_t x, r;
            std::memcpy(&x, &a[i], 4);
            r = (*f)(x);
            std::memcpy(&out[i], &r, 4);
        }
    }
    return 0;
}

// floating point unary / binary / ternary op in precision prec (4 = float
// overloads, 8 = double); values travel as double.
int chai_ref_fun1_fp(int code, int prec, const double* a, double* out, size_t n)
{
    const SimpleFun1* f = fun1(code);
    if (!f) return -1;
    for (size_t i = 0; i < n; i++) out[i] = prec == 4 ? (double)(*f)((float)a[i]) : (*f)(a[i]);
    return 0;
}

int chai_ref_fun2_fp(int code, int prec, const double* a, const double* b, double* out, size_t n)
{
    const SimpleFun2* f = fun2(code);
    if (!f) return -1;
    for (size_t i = 0; i < n; i++)
        out[i] = prec == 4 ? (double)(*f)((float)a[i], (float)b[i]) : (*f)(a[i], b[i]);
    return 0;
}

// ternary op: prec 0 / 1 on uint32 bit patterns, 4 / 8 on doubles
int chai_ref_fun3(int code, int prec, const void* a, const void* b, const void* c, void* out, size_t n)
{
    const SimpleFun3* f = fun3(code);
    if (!f) return -1;
    for (size_t i = 0; i < n; i++) {
        if (prec == 0) {
            ((uint32_t*)out)[i] = (*f)(((const uint32_t*)a)[i], ((const uint32_t*)b)[i], ((const uint32_t*)c)[i]);
        } else if (prec == 1) {
            int32_t x, y, z, r;
            std::memcpy(&x, (const uint32_t*)a + i, 4);
            std::memcpy(&y, (const uint32_t*)b + i, 4);
            std::memcpy(&z, (const uint32_t*)c + i, 4);
            r = (*f)(x, y, z);
            std::memcpy((uint32_t*)out + i, &r, 4);
        } else if (prec == 4) {
            ((double*)out)[i] = (double)(*f)((float)((const double*)a)[i], (float)((const double*)b)[i],
                                             (float)((const double*)c)[i]);
        } else {
            ((double*)out)[i] = (*f)(((const double*)a)[i], ((const double*)b)[i], ((const double*)c)[i]);
        }
    }
    return 0;
}

}  // extern "C"
