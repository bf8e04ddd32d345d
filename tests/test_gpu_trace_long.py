"""Long fused kernels: repeated statements (a traced loop) are emitted as a loop and
compiled; a kernel still longer than the JIT limit after that (CHAI_TRACE_JIT_MAX, 1024 VM
instructions by default; hipRTC's compile time grows faster than linearly with the
kernel's length) is regrouped into shorter compiled kernels, the intermediate arrays
stored between them; a single statement longer than the limit runs on the trace VM.
All give the step-by-step f32 result bit for bit."""
import numpy as np
import pytest
import torch

from chai_amd import F32
from chai_amd import trace as tr

pytestmark = pytest.mark.gpu


def loop_trace(reps, n, varying=False):
    """C = A; then `reps` times C = C * c_i + B (a traced C++ loop, unrolled); c_i =
    0.999f each time, or a different constant per statement when `varying`"""
    t = tr.Trace()
    A, B = t.input(F32, n), t.input(F32, n)
    C = t.output(F32, n)
    t.assign(C, A + tr.f32(0.0))
    consts = [np.float32(0.999 - (i * 1e-4 if varying else 0.0)) for i in range(reps)]
    for c in consts:
        t.assign(C, C * tr.f32(float(c)) + B)
    return t, A, B, C, consts


@pytest.mark.parametrize("reps,varying", [(8, False), (400, False), (400, True)])
def test_long_kernels_roll_or_split(chai, reps, varying):
    """400 repetitions of one statement roll into a loop and are compiled as one
    kernel; 400 different statements exceed the JIT limit as one kernel and are
    regrouped into several compiled kernels"""
    n = 4096
    tr.set_jit(True)
    t, A, B, C, consts = loop_trace(reps, n, varying)
    prog = t.compile()
    nk = prog.info()[0]
    print("reps %d varying %s: %d kernels" % (reps, varying, nk), flush=True)
    if varying:
        assert nk > 1
    else:
        assert nk == 1
    for k in range(nk):
        src, jitted = prog.source(k)
        assert jitted, src[:300]
    if reps == 400 and not varying:
        assert "for (int it = 0; it < 400;" in prog.source(0)[0]
    rng = np.random.default_rng(reps)
    a = rng.standard_normal(n).astype(np.float32)
    b = rng.standard_normal(n).astype(np.float32)
    c = torch.empty(n, device="cuda")
    prog.run({A: torch.from_numpy(a).cuda(), B: torch.from_numpy(b).cuda(), C: c})
    torch.cuda.synchronize()
    ref = a + np.float32(0.0)
    for k in consts:
        ref = (ref * k).astype(np.float32) + b
    assert np.array_equal(c.cpu().numpy(), ref)


def test_one_statement_past_the_limit_runs_on_the_vm(chai):
    """A single statement whose kernel exceeds the JIT limit cannot be split:
    C = ((A * c_0 + B) * c_1 + B) ... as one expression, 600 distinct c_i"""
    n = 4096
    tr.set_jit(True)
    t = tr.Trace()
    A, B = t.input(F32, n), t.input(F32, n)
    C = t.output(F32, n)
    consts = [np.float32(0.999 - i * 1e-4) for i in range(600)]
    e = A + tr.f32(0.0)
    for c in consts:
        e = e * tr.f32(float(c)) + B
    t.assign(C, e)
    prog = t.compile()
    assert prog.info()[0] == 1
    src, jitted = prog.source(0)
    assert not jitted and "exceed the JIT limit" in src
    rng = np.random.default_rng(3)
    a = rng.standard_normal(n).astype(np.float32)
    b = rng.standard_normal(n).astype(np.float32)
    c = torch.empty(n, device="cuda")
    prog.run({A: torch.from_numpy(a).cuda(), B: torch.from_numpy(b).cuda(), C: c})
    torch.cuda.synchronize()
    ref = a + np.float32(0.0)
    for k in consts:
        ref = (ref * k).astype(np.float32) + b
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
