"""Pins the oracle's trace evaluator (oracle/trace.c).

- MD5 of known messages as a pure uint32 trace equals hashlib's digests
  (the reference's sample/md5.cpp workload): pins +, &, |, ^, ~, rotate.
- Integer scalar ops against the reference's own misc/SimpleFun{1,2}.cpp
  compiled into oracle/_ref.
- Broadcast, promotion and reductions against numpy closed forms.
"""
import ctypes

import numpy as np
import pytest

from chai_amd import F32, F64, I32, U32
from chai_amd.trace import Trace, f32, i32, u32, mean, sum_, cond, maxe, exp, dot
from tests.trace_cases import (MSGS, md5_expected, md5_hex, md5_inputs, md5_trace, mc_trace,
                               random_trace)


def test_md5_trace_known_answers(oracle):
    t, M, outs = md5_trace(len(MSGS))
    ins = md5_inputs(MSGS)
    r = oracle.trace_eval(t, {m.index: ins[k] for k, m in enumerate(M)})
    got = md5_hex([r[o.index].ravel() for o in outs])
    assert got == md5_expected(MSGS)


INT_CODES = [11, 12, 13, 14, 17, 18, 19, 20, 21, 22, 30, 31, 32, 33, 34, 35, 36, 37, 302, 303,
             401, 402, 405, 411, 412, 413, 414, 415]


@pytest.mark.parametrize("prec", [U32, I32])
def test_int_binary_ops_match_reference_simplefun(oracle, prec):
    ref = oracle.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (no reference tree here)")
    rng = np.random.default_rng(1)
    n = 512
    a = rng.integers(0, 2 ** 32, size=n, dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, 2 ** 32, size=n, dtype=np.uint64).astype(np.uint32)
    b[:64] = rng.integers(1, 32, size=64)            # small shift / rotate counts
    a[64:80] = 0x80000000; b[64:80] = 0xffffffff     # INT_MIN / -1 edge
    for code in INT_CODES:
        bb = b.copy()
        if code in (12, 17):
            bb[bb == 0] = 1                          # the reference traps on x / 0
            if prec == I32:
                bb[(a == 0x80000000) & (bb == 0xffffffff)] = 1
        if code in (18, 19, 414):
            bb = bb % 32 if code != 414 else bb % 31 + 1   # defined shift counts only
        expect = np.empty(n, np.uint32)
        rc = ref.chai_ref_fun2_int(code, prec, a.ctypes.data_as(ctypes.c_void_p),
                                   bb.ctypes.data_as(ctypes.c_void_p),
                                   expect.ctypes.data_as(ctypes.c_void_p), n)
        assert rc == 0, code
        t = Trace()
        va = t.input(prec, n); vb = t.input(prec, n)
        from chai_amd.trace import Expr
        o = t.assign(t.output(prec, n), Expr(code, [va, vb]))
        dt = np.uint32 if prec == U32 else np.int32
        r = oracle.trace_eval(t, {va.index: a.view(dt), vb.index: bb.view(dt)})
        assert np.array_equal(r[o.index].ravel().view(np.uint32), expect), code


def test_int_unary_ops_match_reference_simplefun(oracle):
    ref = oracle.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built")
    from chai_amd.trace import Expr
    rng = np.random.default_rng(2)
    a = rng.integers(0, 2 ** 32, size=256, dtype=np.uint64).astype(np.uint32)
    a[0] = 0
    for prec in (U32, I32):
        for code in (15, 16, 400, 10):
            if code == 10 and prec == U32:
                continue
            expect = np.empty(256, np.uint32)
            assert ref.chai_ref_fun1_int(code, prec, a.ctypes.data_as(ctypes.c_void_p),
                                         expect.ctypes.data_as(ctypes.c_void_p), 256) == 0
            t = Trace()
            va = t.input(prec, 256)
            o = t.assign(t.output(prec, 256), Expr(code, [va]))
            dt = np.uint32 if prec == U32 else np.int32
            r = oracle.trace_eval(t, {va.index: a.view(dt)})
            assert np.array_equal(r[o.index].ravel().view(np.uint32), expect), (prec, code)


def test_broadcast_promotion_reduction(oracle):
    """(x % W, y % H, s % slots) broadcast; f32 + f64 literal -> f64; mean / W*H."""
    t = Trace()
    a = t.input(F32, 6, 4, 2)
    row = t.input(F32, 6, 1, 1)           # broadcast down the height and the slots
    col = t.input(I32, 1, 4, 2)           # broadcast across the width
    o = t.assign(t.output(F64, 6, 4, 2), a * row + col + 0.5)      # 0.5 is an f64 literal
    m = t.assign(t.output(F64, 1, 1, 2), mean(o))
    s = t.assign(t.output(F32, 1, 1, 2), sum_(a))
    rng = np.random.default_rng(0)
    A = rng.random((2, 4, 6), dtype=np.float32)
    R = rng.random((1, 1, 6), dtype=np.float32)
    Cl = rng.integers(-5, 5, size=(2, 4, 1)).astype(np.int32)
    r = oracle.trace_eval(t, {a.index: A, row.index: R, col.index: Cl})
    # f32 * f32 -> f32; f32 + i32 -> f32 (int converted to float); + f64 literal -> f64
    expect = ((A * R).astype(np.float32) + Cl.astype(np.float32)).astype(np.float32).astype(np.float64) + 0.5
    assert np.array_equal(r[o.index], expect)
    assert np.allclose(r[m.index].ravel(), expect.reshape(2, -1).mean(1), rtol=1e-15)
    assert np.array_equal(r[s.index].ravel(), A.astype(np.float64).reshape(2, -1).sum(1).astype(np.float32))


def test_cond_and_dot(oracle):
    t = Trace()
    x = t.input(F32, 50)
    y = t.input(I32, 50)
    c = t.assign(t.output(F32, 50), cond(x > f32(0.5), x, y))
    d = t.assign(t.output(F32, 1), dot(x, y))
    rng = np.random.default_rng(3)
    X = rng.random(50, dtype=np.float32)
    Y = rng.integers(-9, 9, size=50).astype(np.int32)
    r = oracle.trace_eval(t, {x.index: X, y.index: Y})
    assert np.array_equal(r[c.index].ravel(), np.where(X > 0.5, X, Y.astype(np.float32)))
    assert r[d.index].ravel()[0] == np.float32((X * Y.astype(np.float32)).astype(np.float64).sum())


def test_mc_trace_small(oracle):
    t, m = mc_trace(1 << 14)
    r = oracle.trace_eval(t, {})
    z = oracle.rng_fill(1230, 4, 4, 7, 1 << 14, normal=True).ravel()
    S = np.float32(100) * np.exp(np.float32(0.03) + np.float32(0.2) * z)
    # numpy's f32 exp and glibc's expf may differ in the last bit; the mean agrees closely
    assert abs(r[m.index].ravel()[0] - np.maximum(S - np.float32(105), 0).astype(np.float64).mean()) < 1e-4


@pytest.mark.parametrize("seed", range(4))
def test_random_traces_run_on_oracle(oracle, seed):
    for kind in ("int", "fp"):
        t, data, outs, red = random_trace(seed, kind)
        r = oracle.trace_eval(t, data)
        assert r[red.index].size == 2


# ------------------------------------------------------------------- gathers
def test_gather_filter_sample_is_a_periodic_stencil(oracle):
    """sample/filter.cpp (N = 12, as the sample): five gather2_floor reads at
    (X +- 1, Y +- 1) equal numpy's periodic rolls, summed in the same order"""
    from tests.trace_cases import filter_expected, filter_trace
    t, A, B = filter_trace(12)
    a = np.random.default_rng(1).standard_normal((1, 12, 12)).astype(np.float32)
    r = oracle.trace_eval(t, {A.index: a})
    assert np.array_equal(r[B.index], filter_expected(a))


def test_gather_lattice_sample_is_a_binomial_step(oracle):
    """sample/finance/american_option.cpp: gather1_floor(option, IDX) and
    (option, IDX + 1) are the option array and its left shift (IDX + 1 = N
    wraps to 0); six backward steps equal numpy's recurrence bit for bit"""
    from tests.trace_cases import LATTICE, lattice_expected, lattice_trace
    N, steps = 101, 6
    t, (P, O), (Pn, On) = lattice_trace(N, steps, **LATTICE)
    p0 = 100.0 * LATTICE["up"] ** (2.0 * np.arange(N) - (N - 1))
    o0 = np.maximum(p0 - LATTICE["strike"], 0.0)
    r = oracle.trace_eval(t, {P.index: p0, O.index: o0})
    ep, eo = lattice_expected(p0, o0, steps, **LATTICE)
    assert np.array_equal(r[Pn.index].ravel(), ep) and np.array_equal(r[On.index].ravel(), eo)


@pytest.mark.parametrize("prec", [U32, I32, F32, F64])
@pytest.mark.parametrize("two", [False, True])
def test_gather_subscript_rule_closed_form(oracle, prec, two):
    """every range of the floored subscript (inside, negative above / below
    -W, past W, NaN / inf / out of int range) on a data array of non power
    of two width whose elements hold their own flat index; subscripts with
    more slots than the data"""
    from tests.trace_cases import gather_rule_trace
    t, data, o, expect = gather_rule_trace(prec, two, seed=int(prec) * 2 + two)
    r = oracle.trace_eval(t, data)
    assert np.array_equal(r[o.index], expect)


def test_gather_subscript_closed_form_cases():
    """the closed form itself on hand-worked cases (W = 12): -1 -> 11,
    -12 -> 0, -13 -> (2^64 - 1) mod 12 = 3, 25 -> 1; INT_MIN for NaN"""
    from tests.trace_cases import subscript_closed_form as f
    assert [f(v, I32, 12) for v in (-1, -12, -13, 25, 0, 11)] == [11, 0, 3, 1, 0, 11]
    assert (2 ** 64 - 1) % 12 == 3
    assert f(np.float32(-0.5), F32, 12) == 11 and f(np.float64(np.nan), F64, 12) == (2 ** 64 - 2 ** 31 + 12) % 12
    assert f(np.uint32(2 ** 32 - 1), U32, 8) == 7        # bits of -1; power of two width: true modulo


def test_gather_broadcast_reduction_and_second_kernel(oracle):
    from tests.trace_cases import gather_bcast_trace
    t, data, outs, expect = gather_bcast_trace()
    r = oracle.trace_eval(t, data)
    assert np.array_equal(r[outs[0].index], expect[0])
    assert np.array_equal(r[outs[1].index].ravel(), expect[1])
    assert np.array_equal(r[outs[2].index], expect[2])


def test_gather_data_must_be_an_array():
    from chai_amd.trace import gather1_floor
    t = Trace()
    a = t.input(F32, 8)
    with pytest.raises(TypeError):
        gather1_floor(2.0, a)


# --------------------------------------------------- reassignment, ones / zeros
@pytest.mark.parametrize("NT,N", [(12, 12), (20, 7)])
def test_kirchhoff_sample_reassigns_and_gathers(oracle, NT, N):
    """sample/kirch.cpp: modl assigned N + 1 times, each read seeing the
    previous assignment; equals the numpy restatement bit for bit"""
    from tests.trace_cases import kirchhoff_expected, kirchhoff_trace
    t, data, modl, consts = kirchhoff_trace(NT, N)
    d = np.array([[1.0 / (1 + i * j) for j in range(NT)] for i in range(N)], dtype=np.float32)
    r = oracle.trace_eval(t, {data.index: d})
    assert np.array_equal(r[modl.index][0], kirchhoff_expected(d, NT, N, consts))


def test_ones_zeros_leaves(oracle):
    """ones / zeros leaves in each type; code 73 (declared for both ones2_i32
    and zeros2_f64) is an f64 zero, as the reference's dispatch table keeps
    the later registration"""
    from chai_amd.trace import Expr
    t = Trace()
    outs = [t.assign(t.output(p, 5, 3), Trace.ones(p, 5, 3) + Trace.ones(p, 5)) for p in (U32, I32, F32, F64)]
    z = t.assign(t.output(F64, 5, 3), Trace.zeros(F32, 5, 3) + Expr(73, W=5, H=3) + f32(2.5))
    r = oracle.trace_eval(t, {})
    for o, p in zip(outs, (U32, I32, F32, F64)):
        # ones(I32, 5, 3) is code 73, an f64 zero: 0 + 1 for that row
        assert (r[o.index] == (1 if p == I32 else 2)).all(), p
    assert (r[z.index] == 2.5).all()
    assert Trace.ones(I32, 5, 3).code == 73 and Trace.zeros(F64, 5, 3).code == 73


def test_reassignment_reads_the_latest_value(oracle):
    t = Trace()
    a = t.input(F32, 16)
    s = t.temp(F32, 16)
    o = t.output(F32, 16)
    t.assign(s, a * f32(2.0))
    t.assign(o, s + f32(1.0))
    t.assign(s, s * s)
    t.assign(o, o + s)
    A = np.arange(16, dtype=np.float32)
    r = oracle.trace_eval(t, {a.index: A})
    assert np.array_equal(r[o.index].ravel(), (A * 2 + 1) + (A * 2) * (A * 2))



# ------------------------------------------------------------ statement splitting
def test_loopsum_sample_closed_form(oracle):
    """sample/loopsum.cpp: A = A + mean(B) three times, B += 1, twice; C = sum(A).
    A ends as A + 3*5 + 3*6 and C = sum(A) + 33 * 100"""
    from tests.trace_cases import loopsum_trace
    t, A, C, An = loopsum_trace()
    a = np.arange(100, dtype=np.float32).reshape(1, 1, 100)
    r = oracle.trace_eval(t, {A.index: a})
    assert np.array_equal(r[An.index], a + np.float32(33))
    assert r[C.index].ravel()[0] == np.float32(a.sum() + 33 * 100)


def test_nested_reduction_and_matmul_in_expressions(oracle):
    """a reduction and a matmul below the root of a statement, a matmul of
    expressions with a transposed operand: closed forms on integer data"""
    from tests.trace_cases import split_trace_case
    t, data, outs, expect = split_trace_case()
    r = oracle.trace_eval(t, data)
    for o, e in zip(outs, expect):
        assert np.array_equal(r[o.index].reshape(e.shape), e), o.index
