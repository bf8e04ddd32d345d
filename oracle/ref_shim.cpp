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

}  // extern "C"
