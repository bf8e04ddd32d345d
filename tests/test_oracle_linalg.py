"""Pins the oracle's GEMM / GEMV / reduction restatement (oracle/linalg.c).

Independent evidence: numpy float64 matmul (a library routine), the all-ones
sentinel C == K (autotune/EvergreenMatmulMM.cpp:800-822), integer closed
forms, and brute force on tiny inputs.
"""
import numpy as np
import pytest


def _unif(o, seed, shape):
    n = int(np.prod(shape))
    W = shape[-1]
    return o.rng_fill(1230, 4, 4, seed, W, n // W, limits=(-1.0, 1.0)).reshape(shape)


def test_f64ref_matches_numpy(oracle):
    A = _unif(oracle, 42, (64, 48)); B = _unif(oracle, 43, (48, 80))
    C = oracle.gemm_f64ref(A, B)
    ref = A.astype(np.float64) @ B.astype(np.float64)
    assert np.allclose(C, ref, rtol=0, atol=1e-13)
    for ta, tb in [(True, False), (False, True), (True, True)]:
        Ax = np.ascontiguousarray(A.T) if ta else A
        Bx = np.ascontiguousarray(B.T) if tb else B
        assert np.allclose(oracle.gemm_f64ref(Ax, Bx, ta, tb, nthreads=2), ref, rtol=0, atol=1e-13)


def test_interp_matmul_semantics(oracle):
    """Interpreter: product rounded in the promoted type (float*float -> float),
    accumulated in double, one rounding to the output (InterpMatmulBase.hpp:49-61)."""
    A = _unif(oracle, 1, (16, 33)); B = _unif(oracle, 2, (33, 24))
    C = oracle.matmul_interp(A, B)
    prods = (A[:, :, None] * B[None, :, :]).astype(np.float32)  # f32 products
    ref = prods.astype(np.float64).sum(axis=1).astype(np.float32)
    # summation order of numpy differs from the serial loop: allow 1 ulp of f32
    assert np.allclose(C, ref, rtol=2e-7, atol=1e-7)
    # general: C = alpha*tmp + beta*C in double
    C0 = _unif(oracle, 3, (16, 24))
    G = oracle.matmul_interp(A, B, C=C0, alpha=0.5, beta=-2.0)
    tmp = C.astype(np.float64)
    assert np.array_equal(G, (0.5 * tmp + -2.0 * C0.astype(np.float64)).astype(np.float32))


def test_sentinel_ones(oracle):
    K = 257
    A = np.ones((8, K), np.float32); B = np.ones((K, 12), np.float32)
    assert np.all(oracle.matmul_interp(A, B) == K)
    assert np.all(oracle.gemm_f64ref(A, B) == K)


def test_mixed_int_float(oracle):
    rng = np.random.default_rng(5)
    A = rng.integers(-128, 128, size=(10, 20)).astype(np.int32)
    B = _unif(oracle, 9, (20, 7))
    C = oracle.matmul_interp(A, B)
    assert C.dtype == np.float32
    ref = (A.astype(np.float32)[:, :, None] * B[None]).astype(np.float32).astype(np.float64).sum(1)
    assert np.allclose(C, ref, rtol=2e-7, atol=1e-6)
    assert np.allclose(oracle.gemm_f64ref(A, B), A.astype(np.float64) @ B.astype(np.float64), atol=1e-12)


def test_matvec(oracle):
    A = _unif(oracle, 11, (40, 24)); x = _unif(oracle, 12, (24,))
    y = oracle.matvec_interp(A, x)
    ref = (A * x[None, :]).astype(np.float32).astype(np.float64).sum(1)
    assert np.allclose(y, ref, rtol=2e-7, atol=1e-7)
    # op(A) = (stored A^T)^T: same product through the transposed layout
    yt = oracle.matvec_interp(np.ascontiguousarray(A.T), x, trans_a=True)
    assert np.array_equal(yt, y)


def test_reduce_closed_forms(oracle):
    n = 100000
    a = np.arange(n, dtype=np.float64).reshape(1, n)
    assert oracle.reduce("sum", a)[0] == n * (n - 1) / 2
    assert oracle.reduce("mean", a)[0] == (n - 1) / 2
    u = np.arange(1, 1001, dtype=np.uint32).reshape(2, 500)
    assert list(oracle.reduce("sum", u)) == [125250, 375250]
    i = -np.ones((3, 7), np.int32)
    assert list(oracle.reduce("sum", i)) == [-7, -7, -7]
    f = np.full((1, 1 << 20), 0.1, np.float32)
    # double accumulation of the f32 value 0.1f, rounded to f32 once
    assert oracle.reduce("sum", f)[0] == np.float32(np.float64(np.float32(0.1)) * (1 << 20))


def test_tea_hash_is_deterministic(oracle):
    h1 = oracle.tea_hash([1, 2, 3]); h2 = oracle.tea_hash([1, 2, 3]); h3 = oracle.tea_hash([1, 2, 4])
    assert h1 == h2 and h1 != h3
