"""GPU parity of chai_hip_gemv against the oracle (interp/InterpMatmulBase.hpp:38-47).

Tolerance: |y - y64| <= 2 * N * 2^-24 * sum_k |a_k x_k| (FP32 lane partials,
FP64 across lanes), and the interpreter restatement within the same bound
plus its own rounding."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

U = 2.0 ** -24


def unif(oracle, seed, rows, cols, prec=4):
    return oracle.rng_fill(1230, prec, 0, seed, cols, rows, 1, limits=(-1.0, 1.0)).reshape(rows, cols)


@pytest.mark.parametrize("M,N", [(1, 1), (7, 13), (300, 500), (1024, 1024), (513, 4099)])
@pytest.mark.parametrize("trans", [False, True])
def test_gemv_f32(chai, oracle, M, N, trans):
    A = unif(oracle, 42, M, N)
    xv = unif(oracle, 43, 1, M if trans else N).ravel()
    y = chai.gemv(torch.from_numpy(A).cuda(), torch.from_numpy(xv).cuda(), trans_a=trans).cpu().numpy()
    A64 = A.astype(np.float64).T if trans else A.astype(np.float64)
    ref = A64 @ xv.astype(np.float64)
    absp = np.abs(A64) @ np.abs(xv.astype(np.float64))
    K = A64.shape[1]
    assert np.all(np.abs(y - ref) <= 2 * K * U * absp + 1e-30)
    o = oracle.matvec_interp(A, xv, trans_a=trans)
    assert np.all(np.abs(y - o) <= 3 * K * U * absp + U * np.abs(o) + 1e-30)


def test_gemv_general_f64(chai, oracle):
    A = unif(oracle, 1, 200, 300, prec=8)
    xv = unif(oracle, 2, 1, 300, prec=8).ravel()
    y0 = unif(oracle, 3, 1, 200, prec=8).ravel()
    yt = torch.from_numpy(y0.copy()).cuda()
    chai.gemv(torch.from_numpy(A).cuda(), torch.from_numpy(xv).cuda(), y=yt, general=True, alpha=2.0, beta=-0.5)
    ref = 2.0 * (A @ xv) - 0.5 * y0
    assert np.allclose(yt.cpu().numpy(), ref, rtol=0, atol=1e-12)


def test_gemv_16384_headline(chai, oracle):
    """The 16384 x 16384 f32 GEMV of the Monte Carlo configuration, sampled rows vs FP64."""
    n = 16384
    A = torch.empty(n, n, device="cuda"); x = torch.empty(n, device="cuda")
    chai.rng_fill(A, seed=42, limits=(-1.0, 1.0), W=n, H=n)
    chai.rng_fill(x, seed=42 + n * n, limits=(-1.0, 1.0), W=n, H=1)
    y = chai.gemv(A, x).cpu().numpy()
    rows = np.arange(0, n, 1021)
    Ah = A[torch.from_numpy(rows).cuda()].double().cpu().numpy()
    xh = x.double().cpu().numpy()
    ref = Ah @ xh
    assert np.all(np.abs(y[rows] - ref) <= 2 * n * U * (np.abs(Ah) @ np.abs(xh)))


@pytest.mark.parametrize("M,N", [(5, 70001), (300, 500), (4100, 1030), (1, 4099)])
@pytest.mark.parametrize("trans", [False, True])
def test_gemv_variants_match_the_interpreter(chai, oracle, M, N, trans):
    """Every GEMV tile variant against the interpreter's matvec (each product
    rounded in f32, summed in double): the FP64 sums differ only in order, so
    |y - y_interp| <= 2^-24 |y_interp| + K 2^-52 sum|a x|."""
    A = unif(oracle, 7, M, N)
    xv = unif(oracle, 8, 1, M if trans else N).ravel()
    At, xt = torch.from_numpy(A).cuda(), torch.from_numpy(xv).cuda()
    o = oracle.matvec_interp(A, xv, trans_a=trans).astype(np.float64)
    A64 = A.astype(np.float64).T if trans else A.astype(np.float64)
    absp = np.abs(A64) @ np.abs(xv.astype(np.float64))
    K = A64.shape[1]
    nv = len(chai.gemm_variants(chai.GEMM_GEMV))
    assert nv == 9
    for v in range(nv):
        y = chai.gemm(At, xt.view(-1, 1), trans_a=trans, mode=chai.GEMM_GEMV, variant=v).cpu().numpy().ravel()
        assert np.all(np.abs(y - o) <= U * np.abs(o) + K * 2.0 ** -52 * absp + 1e-30), v


def test_gemv_tune_checks_each_variant(chai, oracle):
    """The tuner's search over the GEMV family: every variant passes its own
    FP64 check (they order the sums differently) and the winner is used."""
    A = torch.from_numpy(unif(oracle, 9, 2000, 3000)).cuda()
    x = torch.from_numpy(unif(oracle, 10, 1, 3000).ravel()).cuda()
    y = torch.empty(2000, 1, device="cuda")
    v, ms = chai.gemm_tune(A, x.view(-1, 1), y, mode=chai.GEMM_GEMV, trials=2)
    assert 0 <= v < 9 and ms > 0
    y1 = chai.gemv(A, x)
    y2 = chai.gemm(A, x.view(-1, 1), mode=chai.GEMM_GEMV, variant=v).view(-1)
    assert torch.equal(y1, y2)
