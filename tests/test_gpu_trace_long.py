"""Long fused kernels: a kernel longer than the JIT limit (CHAI_TRACE_JIT_MAX, 1024 VM
instructions by default) runs on the trace VM instead of being compiled by hipRTC, whose
compile time grows faster than linearly with the kernel's length; a short kernel is
compiled.  Both give the step-by-step f32 result bit for bit."""
import numpy as np
import pytest
import torch

from chai_amd import F32
from chai_amd import trace as tr

pytestmark = pytest.mark.gpu


def loop_trace(reps, n):
    """C = A; then `reps` times C = C * 0.999f + B (a traced C++ loop, unrolled)"""
    t = tr.Trace()
    A, B = t.input(F32, n), t.input(F32, n)
    C = t.output(F32, n)
    t.assign(C, A + tr.f32(0.0))
    for _ in range(reps):
        t.assign(C, C * tr.f32(0.999) + B)
    return t, A, B, C


@pytest.mark.parametrize("reps", [8, 400])
def test_long_kernel_runs_on_the_vm(chai, reps):
    n = 4096
    tr.set_jit(True)
    t, A, B, C = loop_trace(reps, n)
    prog = t.compile()
    assert prog.info()[0] == 1
    src, jitted = prog.source(0)
    assert jitted == (reps == 8)
    if reps == 400:
        assert "exceed the JIT limit" in src
    rng = np.random.default_rng(reps)
    a = rng.standard_normal(n).astype(np.float32)
    b = rng.standard_normal(n).astype(np.float32)
    c = torch.empty(n, device="cuda")
    prog.run({A: torch.from_numpy(a).cuda(), B: torch.from_numpy(b).cuda(), C: c})
    torch.cuda.synchronize()
    ref = a + np.float32(0.0)
    for _ in range(reps):
        ref = (ref * np.float32(0.999)).astype(np.float32) + b
    assert np.array_equal(c.cpu().numpy(), ref)


def test_continuation_across_a_readout(chai):
    """A readout in the middle of a Chai program splits its trace (the reference
    continues in a new trace with the arrays kept on the device).  Here the first
    trace's output buffer is the second trace's input: the readout in between sees
    the intermediate value, and the end result equals the unsplit trace's."""
    n = 1 << 16
    a = np.random.default_rng(5).standard_normal(n).astype(np.float32)
    ta = torch.from_numpy(a).cuda()
    # whole program as one trace: X = A * 2 + 1; Y = exp(X) * 0.5
    t = tr.Trace()
    A = t.input(F32, n)
    X, Y = t.output(F32, n), t.output(F32, n)
    t.assign(X, A * tr.f32(2.0) + tr.f32(1.0))
    t.assign(Y, tr.exp(X) * tr.f32(0.5))
    x_all, y_all = torch.empty(n, device="cuda"), torch.empty(n, device="cuda")
    t.compile().run({A: ta, X: x_all, Y: y_all})
    # split at a readout of X
    t1 = tr.Trace()
    A1 = t1.input(F32, n)
    X1 = t1.output(F32, n)
    t1.assign(X1, A1 * tr.f32(2.0) + tr.f32(1.0))
    x_dev = torch.empty(n, device="cuda")
    t1.compile().run({A1: ta, X1: x_dev})
    x_read = x_dev.cpu().numpy()                       # the readout
    t2 = tr.Trace()
    X2 = t2.input(F32, n)
    Y2 = t2.output(F32, n)
    t2.assign(Y2, tr.exp(X2) * tr.f32(0.5))
    y_dev = torch.empty(n, device="cuda")
    t2.compile().run({X2: x_dev, Y2: y_dev})
    torch.cuda.synchronize()
    assert np.array_equal(x_read, a * np.float32(2.0) + np.float32(1.0))
    assert torch.equal(x_dev, x_all) and torch.equal(y_dev, y_all)
