"""Inline OpenCL C kernels (ProgramCL, api/ProgramCL.cpp) on the GPU: the
reference samples' computations, eager and inside traces, against closed
forms and numpy restatements (bit-exact for integer and IEEE-exact float
work; 1e-12 relative where OpenCL's exp / pown run).  The first test writes
the work-item functions to clamped indices, so a wrong work-space mapping
fails it without any out-of-bounds access; -x then stops the file."""
import numpy as np
import pytest
import torch

from chai_amd import F32, U32, ChaiError
from chai_amd import program_cl as pcl
from chai_amd import trace as tr
from tests import cl_kernels as K

pytestmark = pytest.mark.gpu


def test_work_item_functions(chai):
    p = pcl.ProgramCL(K.PROBE)
    n, local = 4096, 64
    out = torch.full((4 * n,), 0xFFFFFFFF, dtype=torch.int64).to(torch.int32).cuda()
    p("probe", out, n)(n, local)
    torch.cuda.synchronize()
    o = out.cpu().numpy().view(np.uint32).reshape(n, 4)
    g = np.arange(n, dtype=np.uint32)
    assert np.array_equal(o[:, 0], g)
    assert np.array_equal(o[:, 1], g % local)
    assert np.array_equal(o[:, 2], g // local)
    assert np.all(o[:, 3] == local + 1000 * n)


def test_inline_cl_sample_trace(chai):
    """sample/inline_cl.cpp: C = A + B; (progCL, "add_scalar", C, 1.5f)(N, 1); C += A - 3.3f"""
    n = 20
    A = np.arange(n, dtype=np.float32)
    B = (n - np.arange(n)).astype(np.float32)
    p = pcl.ProgramCL(K.ADD_SCALAR)
    t = tr.Trace()
    a, b = t.input(F32, n), t.input(F32, n)
    c = t.output(F32, n)
    t.assign(c, a + b)
    p("add_scalar", c, 1.5)(n, 1)
    t.assign(c, c + (a - tr.f32(3.3)))
    prog = t.compile()
    assert prog.info()[0] == 3
    ta, tb, tc = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), torch.empty(n, device="cuda")
    prog.run({a: ta, b: tb, c: tc})
    torch.cuda.synchronize()
    ref = (A + B + np.float32(1.5)) + (A - np.float32(3.3))
    assert np.array_equal(tc.cpu().numpy(), ref)


def test_eager_launch_large(chai):
    n = 1 << 22
    x = torch.arange(n, dtype=torch.float32, device="cuda")
    pcl.ProgramCL(K.ADD_SCALAR)("add_scalar", x, 0.25)(n, 256)
    torch.cuda.synchronize()
    assert torch.equal(x, torch.arange(n, dtype=torch.float32, device="cuda") + 0.25)


def test_inline_kernel_on_temp_and_versions(chai):
    """a TEMP mutated in place, then read; a reassigned variable around the kernel"""
    n = 4096
    p = pcl.ProgramCL(K.ADD_SCALAR)
    t = tr.Trace()
    a = t.input(F32, n)
    tmp = t.temp(F32, n)
    d = t.output(F32, n)
    c = t.output(F32, n)
    t.assign(tmp, a * tr.f32(2.0))
    p("add_scalar", tmp, 1.0)(n, 64)
    t.assign(d, tmp + tr.f32(1.0))
    t.assign(c, a + tr.f32(1.0))
    p("add_scalar", c, 0.5)(n, 64)
    t.assign(c, c * tr.f32(3.0))
    A = np.random.default_rng(1).standard_normal(n).astype(np.float32)
    td, tc = torch.empty(n, device="cuda"), torch.empty(n, device="cuda")
    t.compile().run({a: torch.from_numpy(A).cuda(), d: td, c: tc})
    torch.cuda.synchronize()
    assert np.array_equal(td.cpu().numpy(), (A * np.float32(2) + np.float32(1)) + np.float32(1))
    assert np.array_equal(tc.cpu().numpy(), ((A + np.float32(1)) + np.float32(0.5)) * np.float32(3))


@pytest.mark.parametrize("n", [16, 1 << 16])
def test_bitonic_sort_trace(chai, n):
    """sample/sorting/bitonic.cpp: every pass of every stage is an inline
    kernel statement on one array; the result is the sorted keys"""
    keys = np.random.default_rng(n).integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    p = pcl.ProgramCL(K.BITONIC)
    t = tr.Trace()
    a = t.input(U32, n)
    stages = n.bit_length() - 1
    for stage in range(stages):
        for pas in range(stage + 1):
            p("bitonic_pass", a, stage, pas, n, 1)(n // 2, 1 if n < 64 else 64)
    prog = t.compile()
    assert prog.info()[0] == stages * (stages + 1) // 2
    buf = torch.from_numpy(keys.view(np.int32).copy()).cuda()
    prog.run({a: buf})
    torch.cuda.synchronize()
    assert np.array_equal(buf.cpu().numpy().view(np.uint32), np.sort(keys))


def american_ref(strike, vol, s0, r, T, steps):
    """the binomial recurrence of the AMERICAN kernel, in the same order, in double"""
    out = np.empty(len(strike))
    for g in range(len(strike)):
        dt = T / steps
        growth = np.exp(r * dt)
        u = np.exp(vol[g] * np.sqrt(dt))
        d = 1.0 / u
        p = (growth - d) / (u - d)
        s = s0 * d ** steps
        px, opt = np.empty(steps + 1), np.empty(steps + 1)
        for i in range(steps + 1):
            px[i] = s
            opt[i] = max(s - strike[g], 0.0)
            s = s * (u * u)
        for n in range(steps, 0, -1):
            cont = (p * opt[1:n + 1] + (1.0 - p) * opt[:n]) / growth
            px[:n] = d * px[1:n + 1]
            opt[:n] = np.maximum(px[:n] - strike[g], cont)
        out[g] = opt[0]
    return out


def test_american_option_batch(chai):
    """sample/finance/american_option_batch.cpp's sizes: 20 options, 100 steps,
    work groups of 5 with static __local trees, in double"""
    num, steps, group = 20, 100, 5
    strike = np.array([100.0 - num // 2 + i for i in range(num)])
    vol = np.array([0.25 + 0.1 * 4 * (i - num // 2) ** 2 / (num * num) for i in range(num)])
    p = pcl.ProgramCL(K.AMERICAN, "-DSTEPS=%d -DGROUP=%d" % (steps, group))
    val = torch.empty(num, dtype=torch.float64, device="cuda")
    p("american_call", val, torch.from_numpy(strike).cuda(), torch.from_numpy(vol).cuda(), 100.0, 0.10, 1.0)(num, group)
    torch.cuda.synchronize()
    ref = american_ref(strike, vol, 100.0, 0.10, 1.0, steps)
    np.testing.assert_allclose(val.cpu().numpy(), ref, rtol=1e-12, atol=0)
    # an American call on a non-dividend stock is worth its European value: Black-Scholes within the tree's error
    from math import erf, log, sqrt, exp
    N = lambda x: 0.5 * (1 + erf(x / sqrt(2)))
    for k, v, c in zip(strike, vol, val.cpu().numpy()):
        d1 = (log(100.0 / k) + (0.10 + v * v / 2)) / v
        bs = 100.0 * N(d1) - k * exp(-0.10) * N(d1 - v)
        assert abs(c - bs) < 0.05, (k, v, c, bs)


def test_local_pointer_argument(chai):
    n, local = 5000, 64
    groups = (n + local - 1) // local
    x = np.random.default_rng(3).standard_normal(n).astype(np.float32)
    out = torch.empty(groups, device="cuda")
    p = pcl.ProgramCL(K.BLOCK_SUM)
    p("block_sum", out, torch.from_numpy(x).cuda(), local, n)(groups * local, local)
    torch.cuda.synchronize()
    xp = np.zeros(groups * local, np.float32)
    xp[:n] = x
    xp = xp.reshape(groups, local)
    s = np.zeros(groups, np.float32)
    for i in range(local):
        s = (s + xp[:, i]).astype(np.float32)
    assert np.array_equal(out.cpu().numpy(), s)


def test_two_dimensional_work_space(chai):
    W, H = 96, 40
    src = np.random.default_rng(4).standard_normal((W, H)).astype(np.float32)      # in[x * H + y]
    out = torch.empty(H, W, device="cuda")
    pcl.ProgramCL(K.SCALE_T)("scale_t", out, torch.from_numpy(src).cuda(), 0.5, W, H)(W, H, 32, 8)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), src.T * np.float32(0.5))


def test_bad_launch_is_refused(chai):
    p = pcl.ProgramCL(K.ADD_SCALAR)
    x = torch.zeros(20, device="cuda")
    with pytest.raises(ChaiError, match="takes 2 arguments"):
        p("add_scalar", x)(20, 1)
    with pytest.raises(ChaiError, match="not a multiple"):
        p("add_scalar", x, 1.0)(20, 8)
    torch.cuda.synchronize()
    assert torch.count_nonzero(x).item() == 0
