"""CPU check of the fused-trace code generator: the straight-line HIP source
the JIT emits for a trace (chai_hip_trace_emit_source) is compiled for the host
with clang++ against a small shim (tests/emu_shim.hpp), every (block, thread)
of the launch is run sequentially, and the stored outputs are compared with
the oracle's interpreter-semantics evaluation.  The VM kernel runs the same
instruction stream through the same device functions, so this pins the
compiler's lowering (operand order, conversions, register reuse, nested
broadcast chains) without a GPU."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from chai_amd import F32, F64, I32, U32

from tests.trace_cases import (bcast_trace, filter_expected, filter_trace, gather_bcast_trace, gather_rule_trace,
                               md5_expected, md5_hex, md5_inputs, md5_trace, MSGS, random_trace)

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "chai_amd", "csrc")
CLANG = "/opt/rocm/llvm/bin/clang++"

pytestmark = pytest.mark.skipif(not os.path.exists(CLANG), reason="needs the ROCm clang++ for a host build")

DRIVER = r'''
thread_local EmuDim blockIdx, threadIdx, gridDim, blockDim;
extern "C" void emu_run(void** bufs, const uint32_t* slots, int nslots, unsigned S, unsigned gx, double* partial) {
    chai_hip::VmLaunchArgs L; memset(&L, 0, sizeof L);
    for (int k = 0; k < nslots; k++) { L.vars[k].in = bufs[slots[k]]; L.vars[k].out = bufs[slots[k]]; }
    L.partial = partial;
    gridDim = {gx, S, 1}; blockDim = {256, 1, 1};
    for (unsigned y = 0; y < S; y++) for (unsigned bx = 0; bx < gx; bx++) for (unsigned t = 0; t < 256; t++) {
        blockIdx = {bx, y, 0}; threadIdx = {t, 0, 0}; chai_trace_kernel(L);
    }
}
'''


def emulate(t, data, tmp_path, kernel=0, slots_space=None):
    from chai_amd import trace as tr
    from oracle import oracle as o
    src = tr.emit_source(t, kernel)
    body = src[src.index('#include "trace_vm.hpp"'):]
    slots = [int(m) for m in re.findall(r"// L.vars\[\d+\] = trace variable (\d+)", body)]
    vps = int(re.search(r"v < (\d+)ull", body).group(1))
    tmp_path.mkdir(parents=True, exist_ok=True)
    cpp, so = tmp_path / "k.cpp", tmp_path / "k.so"
    cpp.write_text(body.replace('extern "C" __global__', 'extern "C"') + DRIVER)
    subprocess.run([CLANG, "-O1", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", "-w",
                    "-include", os.path.join(HERE, "emu_shim.hpp"), "-I", CSRC, str(cpp), "-o", str(so)],
                   check=True)
    lib = ctypes.CDLL(str(so))
    arrs = {}
    for i, (p, role, W, H, S) in enumerate(t.vars):
        dt = o.PREC_DTYPE[p]
        arrs[i] = (np.ascontiguousarray(data[i], dtype=dt).reshape(S, H, W).copy() if i in data
                   else np.zeros((S, H, W), dtype=dt))
    bufs = (ctypes.c_void_p * len(arrs))(*[arrs[i].ctypes.data for i in range(len(arrs))])
    sl = (ctypes.c_uint32 * len(slots))(*slots)
    S = slots_space or max(v[4] for v in t.vars)     # slots of the kernel's index space
    part = np.zeros(1 << 16)
    lib.emu_run(bufs, sl, len(slots), S, max(1, (vps + 255) // 256), part.ctypes.data_as(ctypes.c_void_p))
    return arrs, o.trace_eval(t, data)


@pytest.mark.parametrize("seed", range(16))
def test_emitted_bcast_traces_match_oracle(seed, tmp_path):
    t, data, outs = bcast_trace(seed)
    got, ref = emulate(t, data, tmp_path)
    for v in outs:
        np.testing.assert_array_equal(got[v.index], ref[v.index], err_msg="seed %d var %d" % (seed, v.index))


@pytest.mark.parametrize("kind", ["int", "fp"])
@pytest.mark.parametrize("seed", [0, 3, 8, 11])
def test_emitted_random_traces_match_oracle(kind, seed, tmp_path):
    t, data, outs, red = random_trace(seed, kind)
    got, ref = emulate(t, data, tmp_path)
    for v in outs:
        np.testing.assert_array_equal(got[v.index], ref[v.index], err_msg="%s seed %d var %d" % (kind, seed, v.index))


@pytest.mark.parametrize("N", [12, 37])
def test_emitted_filter_stencil_matches_oracle(N, tmp_path):
    """sample/filter.cpp's gather2_floor stencil through the emitted kernel"""
    t, A, B = filter_trace(N)
    a = np.random.default_rng(N).standard_normal((1, N, N)).astype(np.float32)
    got, ref = emulate(t, {A.index: a}, tmp_path)
    np.testing.assert_array_equal(got[B.index], ref[B.index])
    np.testing.assert_array_equal(got[B.index], filter_expected(a))


@pytest.mark.parametrize("prec", [U32, I32, F32, F64])
@pytest.mark.parametrize("two", [False, True])
def test_emitted_gather_subscripts_match_oracle(prec, two, tmp_path):
    """every range of the subscript rule, subscript slots broadcast by modulo"""
    t, data, o, expect = gather_rule_trace(prec, two, seed=int(prec) * 2 + two)
    got, ref = emulate(t, data, tmp_path, slots_space=2)
    np.testing.assert_array_equal(got[o.index], ref[o.index])
    np.testing.assert_array_equal(got[o.index], expect)


def test_emitted_gather_under_broadcast_matches_oracle(tmp_path):
    """gathers smaller than the statement (nested modulo chains) and
    subscript expressions mixing inputs with index arrays (first kernel)"""
    t, data, outs, expect = gather_bcast_trace()
    got, ref = emulate(t, data, tmp_path, kernel=0, slots_space=4)
    np.testing.assert_array_equal(got[outs[0].index], ref[outs[0].index])


def test_emitted_kirchhoff_reassignment_matches_oracle(tmp_path):
    """sample/kirch.cpp: thirteen assignments of modl fused into one kernel"""
    from tests.trace_cases import kirchhoff_expected, kirchhoff_trace
    t, data, modl, consts = kirchhoff_trace(12, 12)
    d = np.array([[1.0 / (1 + i * j) for j in range(12)] for i in range(12)], dtype=np.float32)
    got, ref = emulate(t, {data.index: d}, tmp_path)
    np.testing.assert_array_equal(got[modl.index], ref[modl.index])
    np.testing.assert_array_equal(got[modl.index][0], kirchhoff_expected(d, 12, 12, consts))


def test_emitted_md5_matches_hashlib(tmp_path):
    t, M, outs = md5_trace(len(MSGS))
    data = {m.index: w for m, w in zip(M, md5_inputs(MSGS))}
    got, ref = emulate(t, data, tmp_path)
    digests = md5_hex([got[v.index].reshape(-1) for v in outs])
    assert digests == md5_expected(MSGS)


@pytest.mark.parametrize("vl", [0, 1, 2])
def test_emitted_rng_normals_match_oracle(vl, tmp_path):
    """f32 and f64 Box-Muller normals of the device code (host build) equal the
    oracle's stream bit for bit, including vector lengths 1 and 2."""
    from chai_amd import F32, F64
    from chai_amd.trace import Trace
    t = Trace()
    a = t.assign(t.output(F32, 1001, 3, 2), Trace.rng_normal(F32, 1001, 3, 2, seed=7, vec_len=vl))
    b = t.assign(t.output(F64, 1001, 3, 2), Trace.rng_normal(F64, 1001, 3, 2, seed=9, vec_len=min(vl, 2) or 0))
    got, ref = emulate(t, {}, tmp_path)
    np.testing.assert_array_equal(got[a.index], ref[a.index])
    np.testing.assert_array_equal(got[b.index], ref[b.index])


def _ops_trace(codes, prec, arity, n, seed):
    """one statement out_k = op_k(x0[, x1[, x2]]) per op code over shared inputs
    of n mixed values (at most 20 ops, so that they fuse into one kernel)"""
    from chai_amd.trace import Expr
    from tests.test_oracle_ops import fp_values, int_values
    t = Trace()
    ins, data = [], {}
    for k in range(arity):
        v = t.input(prec, n)
        vals = fp_values(prec, n, seed + k) if prec in (F32, F64) else int_values(prec, n, seed + k)
        data[v.index] = np.roll(vals, 3 * k)[:n]
        ins.append(v)
    outs = [t.assign(t.output(prec, n), Expr(code, ins)) for code in codes]
    return t, data, outs


from chai_amd.trace import Trace  # noqa: E402


def _check_ops(prec, groups, tmp_path):
    from tests.test_oracle_ops import same_bits
    bad = []
    for arity, codes in groups:
        for i in range(0, len(codes), 20):
            chunk = codes[i:i + 20]
            t, data, outs = _ops_trace(chunk, prec, arity, 300, 100 * arity + i)
            got, ref = emulate(t, data, tmp_path / ("a%d_%d" % (arity, i)))
            bad += [c for c, o in zip(chunk, outs) if not same_bits(got[o.index].ravel(), ref[o.index].ravel())]
    assert not bad, bad


@pytest.mark.parametrize("prec", [F32, F64])
def test_emitted_scalar_fp_ops_match_oracle(prec, tmp_path):
    """every floating point op of the VM (device semantics built for the host,
    which calls the same libm as the oracle) equals the oracle bit for bit"""
    from tests.test_oracle_ops import FP1, FP2, FP3
    _check_ops(prec, ((1, FP1), (2, FP2), (3, FP3)), tmp_path)


@pytest.mark.parametrize("prec", [U32, I32])
def test_emitted_scalar_int_ops_match_oracle(prec, tmp_path):
    from tests.test_oracle_ops import INT3
    int2 = [11, 12, 13, 14, 17, 18, 19, 20, 21, 22, 30, 31, 32, 33, 34, 35, 36, 37, 302, 303, 401, 402, 405,
            411, 412, 413, 414, 415]
    int1 = [10, 15, 16, 400, 404, 225]
    _check_ops(prec, ((1, int1), (2, int2), (3, INT3)), tmp_path)


def _loop_trace(reps, n=37):
    from chai_amd.trace import Trace, f32
    t = Trace()
    A, B = t.input(F32, n, 3), t.input(F32, n, 3)
    C = t.output(F32, n, 3)
    t.assign(C, A + f32(0.0))
    for _ in range(reps):
        t.assign(C, C * f32(0.999) + B)
    return t, A, B, C


@pytest.mark.parametrize("reps", [3, 20, 21])
def test_rolled_loop_matches_oracle(reps, tmp_path):
    """a traced loop (one carried variable) is emitted as a for loop with a
    register shuffle, and computes what the unrolled statements compute"""
    from chai_amd import trace as tr
    t, A, B, C = _loop_trace(reps)
    src = tr.emit_source(t, 0)
    assert ("for (int it = 0; it < %d;" % reps) in src or reps < 4    # 5-instruction body, phase-shifted
    rng = np.random.default_rng(reps)
    data = {A.index: rng.standard_normal((1, 3, 37)).astype(np.float32),
            B.index: rng.standard_normal((1, 3, 37)).astype(np.float32)}
    got, ref = emulate(t, data, tmp_path)
    np.testing.assert_array_equal(got[C.index], ref[C.index])


@pytest.mark.parametrize("reps", [6, 7])
def test_rolled_loop_two_carried_variables(reps, tmp_path):
    """two variables carried through the loop (a rotation), an odd and an even
    trip count (the registers' final positions differ), an integer counter too"""
    from chai_amd.trace import Trace, f32
    from chai_amd import trace as tr
    t = Trace()
    X0, Y0 = t.input(F32, 50), t.input(F32, 50)
    X, Y, K = t.output(F32, 50), t.output(F32, 50), t.output(I32, 50)
    t.assign(X, X0 * f32(1.0))
    t.assign(Y, Y0 * f32(1.0))
    t.assign(K, Trace.index(I32, 0, 50))
    for _ in range(reps):
        t.assign(X, X * f32(0.8) - Y * f32(0.6))
        t.assign(Y, X * f32(0.6) + Y * f32(0.8))
        t.assign(K, K * 3 + 1)
    src = tr.emit_source(t, 0)
    assert "for (int it = 0; it < %d;" % (reps - 1) in src     # the last repetition stores instead
    rng = np.random.default_rng(reps)
    data = {X0.index: rng.standard_normal((1, 1, 50)).astype(np.float32),
            Y0.index: rng.standard_normal((1, 1, 50)).astype(np.float32)}
    got, ref = emulate(t, data, tmp_path)
    for v in (X, Y, K):
        np.testing.assert_array_equal(got[v.index], ref[v.index])


def test_varying_statements_are_not_rolled():
    from chai_amd.trace import Trace, f32
    from chai_amd import trace as tr
    t = Trace()
    A = t.input(F32, 64)
    C = t.output(F32, 64)
    t.assign(C, A + f32(0.0))
    for i in range(12):
        t.assign(C, C * f32(1.0 + i / 64.0) + A)
    assert "for (int it" not in tr.emit_source(t, 0)
