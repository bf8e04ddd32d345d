"""GPU parity of fused traces (chai_hip_trace_*) against the oracle's
interpreter-semantics evaluation.

Bar: integer traces bit-exact; floating point traces built from IEEE
operations (+ - * / comparisons max min, conversions) bit-exact; reductions
(FP64 accumulation in a different order) within n * 2^-52 * sum|x|;
transcendentals (device math library vs glibc) within 4 ulp of the result
per element (float: 4 * 2^-23 relative, plus 1e-30 absolute).
"""
import numpy as np
import pytest
import torch

from chai_amd import F32, F64, I32, U32
from chai_amd.trace import (Trace, exp, f32, log, maxe, matmul, mean, sin, sqrt, sum_, tanh,
                            u32, cos, Expr, gather1_floor, gather2_floor)
from tests.trace_cases import (MSGS, bcast_trace, md5_expected, md5_hex, md5_inputs, md5_trace,
                               mc_trace, random_trace)

pytestmark = pytest.mark.gpu


@pytest.fixture(params=["jit", "vm"])
def engine(request, chai):
    """every fused kernel is checked on both engines: hipRTC JIT and trace VM"""
    from chai_amd.trace import set_jit
    set_jit(request.param == "jit")
    yield request.param
    set_jit(True)

TDT = {U32: torch.int32, I32: torch.int32, F32: torch.float32, F64: torch.float64}
NDT = {U32: np.uint32, I32: np.int32, F32: np.float32, F64: np.float64}


def run_gpu(t, inputs):
    """inputs {index: numpy}; returns {index: numpy} of every INPUT/OUTPUT var."""
    prog = t.compile()
    bufs = {}
    for i, (p, role, W, H, S) in enumerate(t.vars):
        if role == 0:
            a = np.ascontiguousarray(inputs[i], dtype=NDT[p]).reshape(-1)
            bufs[i] = torch.from_numpy(a.view(np.int32) if p == U32 else a).cuda()
        elif role == 1:
            bufs[i] = torch.zeros(S * H * W, dtype=TDT[p], device="cuda")
    prog.run(bufs)
    torch.cuda.synchronize()
    out = {}
    for i, b in bufs.items():
        p, _, W, H, S = t.vars[i]
        a = b.cpu().numpy()
        out[i] = (a.view(np.uint32) if p == U32 else a).reshape(S, H, W)
    return out, prog


def test_md5_known_messages(chai, oracle, engine):
    t, M, outs = md5_trace(len(MSGS))
    ins = md5_inputs(MSGS)
    r, prog = run_gpu(t, {m.index: ins[k] for k, m in enumerate(M)})
    assert md5_hex([r[o.index].ravel() for o in outs]) == md5_expected(MSGS)
    nk, ni = prog.info()
    assert nk == 1, "the whole MD5 trace fuses into one kernel"
    assert prog.source(0)[1] == (engine == "jit"), prog.source(0)[0][-2000:]
    print("md5 trace: kernels", nk, "VM instructions", ni)


def test_md5_many_messages(chai, oracle, engine):
    rng = np.random.default_rng(0)
    msgs = [bytes(rng.integers(0, 256, size=int(rng.integers(0, 56)), dtype=np.uint8)) for _ in range(3001)]
    t, M, outs = md5_trace(len(msgs))
    ins = md5_inputs(msgs)
    r, _ = run_gpu(t, {m.index: ins[k] for k, m in enumerate(M)})
    assert md5_hex([r[o.index].ravel() for o in outs]) == md5_expected(msgs)


@pytest.mark.parametrize("seed", range(24))
def test_random_int_traces_bit_exact(chai, oracle, seed, engine):
    t, data, outs, red = random_trace(seed, "int")
    print("seed", seed, flush=True)
    o = oracle.trace_eval(t, data)
    g, _ = run_gpu(t, data)
    for v in outs + [red]:
        assert np.array_equal(g[v.index], o[v.index]), v.index


@pytest.mark.parametrize("seed", range(16))
def test_nested_broadcast_traces_bit_exact(chai, oracle, seed, engine):
    """intermediate results smaller than the index space: x % d1 % d2 chains"""
    t, data, outs = bcast_trace(seed)
    print("bcast seed", seed, flush=True)
    o = oracle.trace_eval(t, data)
    g, _ = run_gpu(t, data)
    for v in outs:
        assert np.array_equal(g[v.index], o[v.index]), v.index


@pytest.mark.parametrize("seed", range(16))
def test_random_fp_traces(chai, oracle, seed, engine):
    t, data, outs, red = random_trace(seed, "fp")
    o = oracle.trace_eval(t, data)
    g, _ = run_gpu(t, data)
    for v in outs:
        assert np.array_equal(g[v.index], o[v.index]), v.index
    x = o[outs[-1].index].reshape(2, -1)
    tol = x.shape[1] * 2.0 ** -52 * np.abs(x).sum(1) + 1e-300
    assert np.all(np.abs(g[red.index].ravel() - o[red.index].ravel()) <= tol)


def test_transcendentals_within_ulp(chai, oracle, engine):
    t = Trace()
    x = t.input(F32, 4096)
    outs = [t.assign(t.output(F32, 4096), f(x)) for f in (exp, log, sin, cos, sqrt, tanh)]
    X = (np.random.default_rng(1).random(4096) * 8 + 1e-3).astype(np.float32)
    o = oracle.trace_eval(t, {x.index: X})
    g, _ = run_gpu(t, {x.index: X})
    for v in outs:
        a, b = g[v.index].astype(np.float64), o[v.index].astype(np.float64)
        assert np.all(np.abs(a - b) <= 4 * 2.0 ** -23 * np.abs(b) + 1e-30), v.index


def test_mc_trace_full_size(chai, oracle, engine):
    """2^24 Philox4x32-10 normals -> S = 100 exp(0.03 + 0.2 z) -> max(S-105,0) -> mean, one kernel."""
    n = 1 << 24
    t, m = mc_trace(n)
    g, prog = run_gpu(t, {})
    assert prog.info()[0] == 1
    o = oracle.trace_eval(t, {})
    gm, om = float(g[m.index].ravel()[0]), float(o[m.index].ravel()[0])
    # the FP64 sums differ only in order and expf by <= 2 ulp per element: relative 1e-6
    assert abs(gm - om) <= 1e-6 * abs(om), (gm, om)
    # closed form: E[max(S-K,0)] for lognormal S (Black-Scholes, undiscounted)
    from math import erf, exp as mexp, log as mlog, sqrt as msqrt
    N = lambda v: 0.5 * (1 + erf(v / msqrt(2)))
    mu, sig = mlog(100.0) + 0.03, 0.2
    d2 = (mu - mlog(105.0)) / sig
    bs = mexp(mu + sig * sig / 2) * N(d2 + sig) - 105.0 * N(d2)
    assert abs(gm - bs) < 0.02, (gm, bs)
    print("mc mean gpu %.7f oracle %.7f closed form %.7f" % (gm, om, bs))


def test_matmul_with_fused_epilogue(chai, oracle):
    """C = matmul(A, B); y = max(C * 0.5f + 1, 0): GEMM kernel + one VM kernel."""
    n = 96
    t = Trace()
    A = t.input(F32, n, n); B = t.input(F32, n, n)
    C = t.assign(t.temp(F32, n, n), matmul(A, B))
    y = t.assign(t.output(F32, n, n), maxe(C * f32(0.5) + f32(1.0), f32(0.0)))
    rng = np.random.default_rng(2)
    a = (rng.random((n, n)) * 2 - 1).astype(np.float32); b = (rng.random((n, n)) * 2 - 1).astype(np.float32)
    g, prog = run_gpu(t, {A.index: a, B.index: b})
    assert prog.info()[0] == 2
    ref = np.maximum((a.astype(np.float64) @ b.astype(np.float64)) * 0.5 + 1, 0)
    assert np.abs(g[y.index].reshape(n, n) - ref).max() < 1e-4


def test_gemv_in_trace(chai, oracle):
    t = Trace()
    A = t.input(F32, 500, 300); x = t.input(F32, 500)
    y = t.assign(t.output(F32, 300), matmul(A, x))
    rng = np.random.default_rng(3)
    a = (rng.random((300, 500)) * 2 - 1).astype(np.float32); xv = (rng.random(500) * 2 - 1).astype(np.float32)
    g, _ = run_gpu(t, {A.index: a, x.index: xv})
    o = oracle.trace_eval(t, {A.index: a, x.index: xv})
    tol = 500 * 2.0 ** -24 * (np.abs(a) @ np.abs(xv)) * 2
    assert np.all(np.abs(g[y.index].ravel() - o[y.index].ravel()) <= tol)


def test_slots_and_broadcast(chai, oracle, engine):
    t = Trace()
    a = t.input(F32, 33, 7, 3)
    row = t.input(F32, 33, 1, 1)
    s = t.input(I32, 1, 1, 3)
    o = t.assign(t.output(F64, 33, 7, 3), a * row + s + 0.25)
    m = t.assign(t.output(F64, 1, 1, 3), mean(o))
    rng = np.random.default_rng(4)
    d = {a.index: rng.random((3, 7, 33), dtype=np.float32), row.index: rng.random((1, 1, 33), dtype=np.float32),
         s.index: rng.integers(-9, 9, size=(3, 1, 1)).astype(np.int32)}
    g, _ = run_gpu(t, d)
    r = oracle.trace_eval(t, d)
    assert np.array_equal(g[o.index], r[o.index])
    assert np.allclose(g[m.index], r[m.index], rtol=1e-14)


def test_unsupported_trace_is_rejected(chai):
    """a bytecode the back end does not know fails compilation, so the caller
    falls back to the interpreter (runtime/Scheduler.cpp:221-254)"""
    t = Trace()
    x = t.input(F32, 16)
    t.assign(t.output(F32, 16), x - Expr(999, [x]))
    with pytest.raises(chai.ChaiError):
        t.compile()


def test_compile_cache_returns_same_program(chai):
    from chai_amd.trace import cache_stats
    t, m = mc_trace(1000 + 7)                 # a size no other test compiles
    h0, m0, e0 = cache_stats()
    p1 = t.compile()
    h1, m1, e1 = cache_stats()
    assert (h1, m1, e1) == (h0, m0 + 1, e0 + 1)
    p2 = t.compile()
    h2, m2, e2 = cache_stats()
    assert (h2, m2, e2) == (h1 + 1, m1, e1) and p1.info() == p2.info()
    t3, _ = mc_trace(1000 + 9)                # another shape is another program
    t3.compile()
    assert cache_stats()[1:] == (m2 + 1, e2 + 1)


# ops whose result is fixed by IEEE arithmetic (or by an exact definition such
# as fmod / remainder / ldexp / nextafter / ilogb / logb): bit-exact on the GPU
EXACT1 = [10, 223, 261, 256, 225, 212, 253, 255, 266, 233, 241, 246, 301, 305, 308, 507, 508, 509, 510, 513]
EXACT2 = [11, 12, 13, 14, 32, 33, 34, 35, 36, 37, 302, 303, 227, 228, 213, 224, 229, 234, 243, 244, 247, 251,
          306, 500, 501, 502, 503, 504, 505, 506, 511, 512]
EXACT3 = [300, 226, 242, 304, 307]
# device math library vs the host libm: within 8 ulp (relative to max(|r|, 1)
# for results that pass through zero)
LIBM1 = [219, 220, 221, 222, 237, 238, 239, 240, 257, 214, 262, 203, 200, 206, 259, 215, 263, 211, 218, 217,
         201, 202, 204, 205, 208, 209, 216, 235, 260, 264, 265]
LIBM2 = [248, 207, 232, 210, 250, 254, 249]


def _ops_gpu(oracle, prec, arity, codes, seed):
    from chai_amd.trace import Expr
    from tests.test_oracle_ops import fp_values
    n = 700
    t = Trace()
    ins, data = [], {}
    for k in range(arity):
        v = t.input(prec, n)
        data[v.index] = np.roll(fp_values(prec, n, seed + k), 3 * k)[:n]
        ins.append(v)
    outs = [(c, t.assign(t.output(prec, n), Expr(c, ins))) for c in codes]
    return outs, oracle.trace_eval(t, data), run_gpu(t, data)[0]


def same_values(a, b):
    """bit-exact except for the sign / payload of NaN, which differ by platform
    (an invalid operation yields a negative default NaN on the x86 host and a
    positive one on the GPU); NaN positions must agree"""
    a, b = np.asarray(a), np.asarray(b)
    na, nb = np.isnan(a), np.isnan(b)
    if not np.array_equal(na, nb):
        return False
    ua = a[~na].view(np.uint32 if a.dtype == np.float32 else np.uint64)
    ub = b[~nb].view(np.uint32 if b.dtype == np.float32 else np.uint64)
    return np.array_equal(ua, ub)


@pytest.mark.parametrize("prec", [F32, F64])
def test_exact_scalar_ops_bit_exact(chai, oracle, engine, prec):
    same_bits = same_values
    bad = []
    for arity, codes in ((1, EXACT1), (2, EXACT2), (3, EXACT3)):
        for i in range(0, len(codes), 16):
            outs, o, g = _ops_gpu(oracle, prec, arity, codes[i:i + 16], 50 * arity + i)
            bad += [c for c, v in outs if not same_bits(g[v.index].ravel(), o[v.index].ravel())]
    assert not bad, bad


@pytest.mark.parametrize("prec", [F32, F64])
def test_libm_scalar_ops_within_ulp(chai, oracle, engine, prec):
    eps = 2.0 ** -23 if prec == F32 else 2.0 ** -52
    bad = []
    for arity, codes in ((1, LIBM1), (2, LIBM2)):
        for i in range(0, len(codes), 16):
            outs, o, g = _ops_gpu(oracle, prec, arity, codes[i:i + 16], 70 * arity + i)
            for c, v in outs:
                a, b = g[v.index].ravel().astype(np.float64), o[v.index].ravel().astype(np.float64)
                nan_ok = np.isnan(a) == np.isnan(b)
                fin = np.isfinite(b) & np.isfinite(a)
                inf_ok = np.where(np.isinf(b) | np.isinf(a), a == b, True)
                # results near overflow (tgamma, pow) may round to inf on one side only
                big = np.abs(b) > (3e38 if prec == F32 else 1.7e308)
                with np.errstate(invalid="ignore"):     # inf - inf where both are the same infinity
                    close = np.abs(a - b) <= 8 * eps * np.maximum(np.abs(b), 1.0)
                if not (np.all(nan_ok) and np.all(inf_ok | big) and np.all(close | ~fin)):
                    bad.append(c)
    assert not bad, bad


@pytest.mark.parametrize("prec", [U32, I32])
def test_integer_ternary_ops_bit_exact(chai, oracle, engine, prec):
    from chai_amd.trace import Expr
    from tests.test_oracle_ops import int_values
    n = 700
    t = Trace()
    ins = [t.input(prec, n) for _ in range(3)]
    data = {v.index: np.roll(int_values(prec, n, 9 + k), 5 * k)[:n] for k, v in enumerate(ins)}
    outs = [t.assign(t.output(prec, n), Expr(c, ins)) for c in (300, 406, 407, 408)]
    outs.append(t.assign(t.output(prec, n), Expr(10, [ins[0]])))       # negation
    o = oracle.trace_eval(t, data)
    g, _ = run_gpu(t, data)
    for v in outs:
        assert np.array_equal(g[v.index], o[v.index]), v.index


# ------------------------------------------------------------------- gathers
@pytest.mark.parametrize("N", [12, 2048])
def test_gather_filter_stencil_bit_exact(chai, oracle, engine, N):
    """sample/filter.cpp's five gather2_floor reads; N = 12 against the
    oracle, both sizes against numpy's periodic rolls (exact in f32)"""
    from tests.trace_cases import filter_expected, filter_trace
    t, A, B = filter_trace(N)
    a = np.random.default_rng(N).standard_normal((1, N, N)).astype(np.float32)
    r, prog = run_gpu(t, {A.index: a})
    assert np.array_equal(r[B.index], filter_expected(a))
    if N <= 64:
        assert np.array_equal(r[B.index], oracle.trace_eval(t, {A.index: a})[B.index])
    assert prog.info()[0] == 1


def test_gather_lattice_one_kernel_per_step(chai, oracle, engine):
    """american_option.cpp's backward induction: each step gathers the
    previous step's temp, so the trace splits into one kernel per step;
    f64 values bit-exact against the oracle and numpy's recurrence"""
    from tests.trace_cases import LATTICE, lattice_expected, lattice_trace
    N, steps = 101, 6
    t, (P, O), (Pn, On) = lattice_trace(N, steps, **LATTICE)
    p0 = 100.0 * LATTICE["up"] ** (2.0 * np.arange(N) - (N - 1))
    o0 = np.maximum(p0 - LATTICE["strike"], 0.0)
    r, prog = run_gpu(t, {P.index: p0, O.index: o0})
    ref = oracle.trace_eval(t, {P.index: p0, O.index: o0})
    ep, eo = lattice_expected(p0, o0, steps, **LATTICE)
    for v, e in ((Pn, ep), (On, eo)):
        assert np.array_equal(r[v.index], ref[v.index])
        assert np.array_equal(r[v.index].ravel(), e)
    assert prog.info()[0] == steps


@pytest.mark.parametrize("prec", [U32, I32, F32, F64])
@pytest.mark.parametrize("two", [False, True])
def test_gather_subscript_rule_bit_exact(chai, oracle, engine, prec, two):
    """every range of the subscript rule (negative above / below -W, past W,
    NaN, inf, out of int range), subscripts with more slots than the data"""
    from tests.trace_cases import gather_rule_trace
    t, data, o, expect = gather_rule_trace(prec, two, seed=int(prec) * 2 + two)
    r, _ = run_gpu(t, data)
    assert np.array_equal(r[o.index], oracle.trace_eval(t, data)[o.index])
    assert np.array_equal(r[o.index], expect)


def test_gather_broadcast_reduction_and_second_kernel(chai, oracle, engine):
    """gathers smaller than their statement (modulo broadcast), subscript
    expressions, a gather under a sum reduction and a gather from a temp of
    the same trace"""
    from tests.trace_cases import gather_bcast_trace
    t, data, outs, expect = gather_bcast_trace()
    r, prog = run_gpu(t, data)
    ref = oracle.trace_eval(t, data)
    for o, e in zip(outs, expect):
        assert np.array_equal(r[o.index], ref[o.index])
        assert np.array_equal(r[o.index].ravel(), e.ravel())
    assert prog.info()[0] >= 3


# --------------------------------------------------- reassignment, ones / zeros
@pytest.mark.parametrize("NT,N", [(12, 12), (1024, 64)])
def test_kirchhoff_reassignment_bit_exact(chai, oracle, engine, NT, N):
    """sample/kirch.cpp: modl assigned N + 1 times (versions in registers),
    N gather2_floor reads, one kernel on the VM; under the JIT, a kernel
    longer than the JIT limit is regrouped into shorter compiled ones.
    Bit-exact against the oracle and the numpy restatement either way"""
    from tests.trace_cases import kirchhoff_expected, kirchhoff_trace
    t, data, modl, consts = kirchhoff_trace(NT, N)
    d = np.array([[1.0 / (1 + i * j) for j in range(NT)] for i in range(N)], dtype=np.float32)
    r, prog = run_gpu(t, {data.index: d})
    assert np.array_equal(r[modl.index], oracle.trace_eval(t, {data.index: d})[modl.index])
    assert np.array_equal(r[modl.index][0], kirchhoff_expected(d, NT, N, consts))
    nk = prog.info()[0]
    if engine == "vm" or N == 12:
        assert nk == 1
    else:
        assert nk > 1 and all(prog.source(k)[1] for k in range(nk))


def test_reassigned_temp_gathered_across_kernels(chai, oracle, engine):
    """a version of a reassigned temp that a later statement gathers from is
    stored and the gathering statement starts a new kernel; the output is
    assigned twice and read between"""
    t = Trace()
    a = t.input(F32, 64, 2)
    s = t.temp(F32, 64, 2)
    o = t.output(F32, 64, 2)
    t.assign(s, a * f32(2.0))
    t.assign(o, s + f32(1.0))
    t.assign(s, gather1_floor(s, Trace.index(I32, 0, 64, 2) + 1) + s)
    t.assign(s, s * f32(3.0))
    t.assign(o, o + gather2_floor(s, Trace.index(F32, 0, 64, 2) * f32(0.5), Trace.index(I32, 1, 64, 2) + 1))
    A = np.random.default_rng(5).standard_normal((1, 2, 64)).astype(np.float32)
    r, prog = run_gpu(t, {a.index: A})
    assert np.array_equal(r[o.index], oracle.trace_eval(t, {a.index: A})[o.index])
    s1 = A[0] * np.float32(2)
    s2 = (np.roll(s1, -1, axis=1) + s1) * np.float32(3)
    expect = (s1 + np.float32(1)) + s2[::-1][:, np.arange(64) // 2]
    assert np.array_equal(r[o.index][0], expect)
    assert prog.info()[0] == 3


def test_ones_zeros_leaves_bit_exact(chai, oracle, engine):
    t = Trace()
    outs = [t.assign(t.output(p, 5, 3), Trace.ones(p, 5, 3) + Trace.ones(p, 5) + Trace.zeros(p, 5))
            for p in (U32, I32, F32, F64)]
    r, _ = run_gpu(t, {})
    ref = oracle.trace_eval(t, {})
    for o in outs:
        assert np.array_equal(r[o.index], ref[o.index])


def test_read_before_first_assignment_is_rejected(chai):
    t = Trace()
    o = t.output(F32, 8)
    t.assign(o, o + f32(1.0))
    t.assign(o, o * f32(2.0))
    with pytest.raises(Exception):
        t.compile()


# ------------------------------------------------------------ statement splitting
def test_loopsum_sample_bit_exact(chai, oracle, engine):
    """sample/loopsum.cpp: mean(B) inside A = A + mean(B) is split into its own
    statement; A and B reassigned; closed form A + 33, sum(A) + 3300"""
    from tests.trace_cases import loopsum_trace
    t, A, C, An = loopsum_trace()
    a = np.arange(100, dtype=np.float32).reshape(1, 1, 100)
    r, prog = run_gpu(t, {A.index: a})
    ref = oracle.trace_eval(t, {A.index: a})
    assert np.array_equal(r[An.index], ref[An.index]) and np.array_equal(r[C.index], ref[C.index])
    assert np.array_equal(r[An.index], a + np.float32(33))
    assert r[C.index].ravel()[0] == np.float32(a.sum() + 3300)


def test_reduction_matmul_and_gather_data_below_the_root(chai, oracle, engine):
    """mean / sum inside an expression, matmul of expressions (one transposed)
    inside an expression, gather from an expression: split into statements
    of their own; integer data keeps every result exact"""
    from tests.trace_cases import split_trace_case
    t, data, outs, expect = split_trace_case()
    r, prog = run_gpu(t, data)
    ref = oracle.trace_eval(t, data)
    for o, e in zip(outs, expect):
        assert np.array_equal(r[o.index], ref[o.index]), o.index
        assert np.array_equal(r[o.index].reshape(e.shape), e), o.index
