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

from tests.trace_cases import bcast_trace, md5_expected, md5_hex, md5_inputs, md5_trace, MSGS, random_trace

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


def emulate(t, data, tmp_path, kernel=0):
    from chai_amd import trace as tr
    from oracle import oracle as o
    src = tr.emit_source(t, kernel)
    body = src[src.index('#include "trace_vm.hpp"'):]
    slots = [int(m) for m in re.findall(r"// L.vars\[\d+\] = trace variable (\d+)", body)]
    vps = int(re.search(r"v < (\d+)ull", body).group(1))
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
    S = max(v[4] for v in t.vars)
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
