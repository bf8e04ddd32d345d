"""GPU parity: chai_hip_reduce (sum / mean, FP64 cross-block combine) against
the oracle's double-accumulating interpreter restatement (interp/InterpAccum.cpp).

Integer sums are exact (int64 accumulation == the interpreter's double sum
below 2^53).  f32/f64 differ from the interpreter only by the order of FP64
additions: tolerance |gpu - oracle| <= 1 ulp of the result (f32) and
n * 2^-52 * sum|x| (f64)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("op", ["sum", "mean"])
@pytest.mark.parametrize("n,slots", [(1, 1), (7, 3), (1000, 1), (4097, 2), (1 << 20, 1), (3, 1000)])
def test_reduce_f32(chai, oracle, op, n, slots):
    x = oracle.rng_fill(1230, 4, 4, 5, n, 1, slots, limits=(-1.0, 1.0)).reshape(slots, n)
    g, g64 = chai.reduce(torch.from_numpy(x).cuda(), op, slots=slots)
    o = oracle.reduce(op, x)
    g = g.cpu().numpy()
    assert np.all(np.abs(g - o) <= np.spacing(np.abs(o).astype(np.float32)) + 1e-30), (g, o)
    exact = x.astype(np.float64).sum(1) / (n if op == "mean" else 1)
    assert np.allclose(g64.cpu().numpy(), exact, rtol=0, atol=n * 2.0 ** -52 * 2)


@pytest.mark.parametrize("op", ["sum", "mean"])
def test_reduce_int_exact(chai, oracle, op):
    rng = np.random.default_rng(1)
    xi = rng.integers(-2 ** 31, 2 ** 31, size=(3, 10001), dtype=np.int64).astype(np.int32)
    g, _ = chai.reduce(torch.from_numpy(xi).cuda(), op, slots=3)
    assert np.array_equal(g.cpu().numpy(), oracle.reduce(op, xi))
    xu = rng.integers(0, 2 ** 32, size=(2, 5003), dtype=np.uint64).astype(np.uint32)
    t = torch.from_numpy(xu.view(np.int32)).cuda()
    g, _ = chai.reduce(t, op, slots=2, prec=chai.U32)
    assert np.array_equal(g.cpu().numpy().view(np.uint32), oracle.reduce(op, xu))


def test_reduce_f64(chai, oracle):
    x = oracle.rng_fill(1231, 8, 2, 3, 1 << 18, 1, 2, normal=True).reshape(2, -1)
    g, _ = chai.reduce(torch.from_numpy(x).cuda(), "sum", slots=2)
    o = oracle.reduce("sum", x)
    tol = x.shape[1] * 2.0 ** -52 * np.abs(x).sum(1)
    assert np.all(np.abs(g.cpu().numpy() - o) <= tol)


def test_reduce_closed_form(chai):
    n = 1 << 22
    x = torch.arange(n, dtype=torch.float64, device="cuda")
    g, _ = chai.reduce(x, "sum")
    assert g.item() == n * (n - 1) / 2
    ones = torch.ones(n, dtype=torch.float32, device="cuda")
    g, g64 = chai.reduce(ones, "mean")
    assert g.item() == 1.0 and g64.item() == 1.0
