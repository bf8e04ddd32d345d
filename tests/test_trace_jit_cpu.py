"""CPU check of the trace JIT: the source the front end emits for each trace
workload compiles with hipRTC for gfx950 (no GPU needed to compile)."""
import ctypes
import os

import pytest

from chai_amd.trace import emit_source
from tests.trace_cases import bcast_trace, md5_trace, mc_trace, random_trace

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "chai_amd", "csrc")


def hiprtc_compile(src):
    try:
        rtc = ctypes.CDLL("libhiprtc.so")
    except OSError:
        rtc = ctypes.CDLL("/opt/rocm/lib/libhiprtc.so")
    names = ["rng_device.hpp", "reduce_device.hpp", "trace_vm.hpp"]
    hdr = [open(os.path.join(CSRC, n)).read().encode() for n in names]
    H = (ctypes.c_char_p * 3)(*hdr)
    N = (ctypes.c_char_p * 3)(*[n.encode() for n in names])
    prog = ctypes.c_void_p()
    assert rtc.hiprtcCreateProgram(ctypes.byref(prog), src.encode(), b"t.hip", 3, H, N) == 0
    opts = [b"--offload-arch=gfx950", b"-O3", b"-std=c++17", b"-ffp-contract=off",
            b"-Xclang", b"-target-feature", b"-Xclang", b"-packed-fp32-ops"]
    O = (ctypes.c_char_p * len(opts))(*opts)
    rc = rtc.hiprtcCompileProgram(prog, len(opts), O)
    n = ctypes.c_size_t()
    rtc.hiprtcGetProgramLogSize(prog, ctypes.byref(n))
    log = ctypes.create_string_buffer(n.value + 1)
    rtc.hiprtcGetProgramLog(prog, log)
    rtc.hiprtcDestroyProgram(ctypes.byref(prog))
    return rc, log.value.decode(errors="replace")


@pytest.mark.parametrize("name", ["mc", "md5", "rand_int", "rand_fp", "bcast"])
def test_jit_source_compiles(name):
    if name == "mc":
        t, _ = mc_trace(1 << 20)
    elif name == "md5":
        t = md5_trace(64)[0]
    elif name == "bcast":
        t = bcast_trace(13)[0]
    else:
        t = random_trace(3, name[5:])[0]
    src = emit_source(t, 0)
    assert "chai_trace_kernel" in src
    rc, log = hiprtc_compile(src)
    assert rc == 0, log[-3000:]
