"""Every scalar op of the trace language: the oracle's restatement
(oracle/trace.c) against the reference's own misc/SimpleFun{1,2,3}.cpp
compiled into oracle/_ref, element by element, bit for bit (both sides call
the same host libm, so transcendentals compare exactly here too).  Ops that
take floating point only are checked in f32 and f64, integer-only ops in u32
and i32; the values include signed zeros, infinities, NaN, subnormals and the
32-bit extremes."""
import ctypes

import numpy as np
import pytest

from chai_amd import F32, F64, I32, U32
from chai_amd.trace import Expr, Trace

FP1 = [10, 223, 261, 256, 219, 220, 221, 222, 237, 238, 239, 240, 257, 214, 262, 203, 200, 206, 259, 215,
       263, 225, 212, 253, 255, 266, 211, 218, 217, 201, 202, 204, 205, 208, 209, 216, 233, 235, 241, 246,
       260, 264, 265, 301, 305, 308, 507, 508, 509, 510, 513]
FP2 = [11, 12, 13, 14, 32, 33, 34, 35, 36, 37, 302, 303, 227, 228, 248, 207, 229, 232, 213, 224, 210, 234,
       243, 244, 247, 249, 250, 251, 254, 306, 500, 501, 502, 503, 504, 505, 506, 511, 512]
FP3 = [300, 226, 242, 304, 307]
INT3 = [300, 406, 407, 408]
NDT = {U32: np.uint32, I32: np.int32, F32: np.float32, F64: np.float64}


@pytest.fixture(scope="module")
def ref(oracle):
    from oracle import oracle as o
    L = o.ref_lib()
    if L is None or not hasattr(L, "chai_ref_fun3"):
        pytest.skip("oracle/_ref not built (the reference tree is not present here)")
    return L


def fp_values(prec, n, seed):
    rng = np.random.default_rng(seed)
    special = [0.0, -0.0, 1.0, -1.0, 0.5, -0.5, 2.0, 3.0, 1e-40, -1e-40, 1e30, -1e30, np.inf, -np.inf, np.nan,
               0.25, 0.75, 1.5, 100.0, -7.25, 5e-324 if prec == F64 else 1e-45]
    x = np.concatenate([special, rng.uniform(-6, 6, n), rng.uniform(0.9, 3, n // 4),
                        rng.standard_normal(n // 4) * 1e3])
    return x.astype(NDT[prec])


def int_values(prec, n, seed):
    rng = np.random.default_rng(seed)
    edge = [0, 1, 2, 3, 0x7fffffff, 0x80000000, 0xffffffff, 0xfffffffe, 0x00ffffff, 0x01000000, 31, 32, 33]
    x = np.concatenate([np.array(edge, dtype=np.uint64), rng.integers(0, 2 ** 32, n, dtype=np.uint64),
                        rng.integers(0, 1000, n // 4, dtype=np.uint64)]).astype(np.uint32)
    return x if prec == U32 else x.view(np.int32)


def oracle_eval(oracle, prec, code, args):
    t = Trace()
    n = args[0].size
    vs = [t.input(prec, n) for _ in args]
    o = t.assign(t.output(prec, n), Expr(code, vs))
    r = oracle.trace_eval(t, {v.index: a for v, a in zip(vs, args)})
    return r[o.index].ravel()


def same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype.kind == "f":
        return np.array_equal(a, b, equal_nan=True) and np.array_equal(np.signbit(a), np.signbit(b))
    return np.array_equal(a, b)


def p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


@pytest.mark.parametrize("prec", [F32, F64])
@pytest.mark.parametrize("code", FP1)
def test_unary_fp_ops_match_reference(oracle, ref, prec, code):
    x = fp_values(prec, 400, code)
    got = oracle_eval(oracle, prec, code, [x])
    xd = x.astype(np.float64)
    out = np.empty_like(xd)
    assert ref.chai_ref_fun1_fp(code, prec, p(xd), p(out), xd.size) == 0
    assert same_bits(got, out.astype(NDT[prec])), code


@pytest.mark.parametrize("prec", [F32, F64])
@pytest.mark.parametrize("code", FP2)
def test_binary_fp_ops_match_reference(oracle, ref, prec, code):
    a = fp_values(prec, 300, code)
    b = np.roll(fp_values(prec, 300, code + 1), 7)
    got = oracle_eval(oracle, prec, code, [a, b])
    ad, bd = a.astype(np.float64), b.astype(np.float64)
    out = np.empty_like(ad)
    assert ref.chai_ref_fun2_fp(code, prec, p(ad), p(bd), p(out), ad.size) == 0
    assert same_bits(got, out.astype(NDT[prec])), code


@pytest.mark.parametrize("prec", [F32, F64])
@pytest.mark.parametrize("code", FP3)
def test_ternary_fp_ops_match_reference(oracle, ref, prec, code):
    a, b, c = (np.roll(fp_values(prec, 200, code + k), 3 * k) for k in range(3))
    got = oracle_eval(oracle, prec, code, [a, b, c])
    ad, bd, cd = (v.astype(np.float64) for v in (a, b, c))
    out = np.empty_like(ad)
    assert ref.chai_ref_fun3(code, prec, p(ad), p(bd), p(cd), p(out), ad.size) == 0
    assert same_bits(got, out.astype(NDT[prec])), code


@pytest.mark.parametrize("prec", [U32, I32])
@pytest.mark.parametrize("code", INT3)
def test_ternary_int_ops_match_reference(oracle, ref, prec, code):
    a, b, c = (np.roll(int_values(prec, 300, code + k), 5 * k) for k in range(3))
    got = oracle_eval(oracle, prec, code, [a, b, c])
    au, bu, cu = (np.ascontiguousarray(v).view(np.uint32) for v in (a, b, c))
    out = np.empty_like(au)
    assert ref.chai_ref_fun3(code, 0 if prec == U32 else 1, p(au), p(bu), p(cu), p(out), au.size) == 0
    assert same_bits(got.view(np.uint32), out), code
