"""CPU pins of the tile-cost MLP oracle (oracle/mlp.c): the analytic gradient
against central finite differences of the loss, and the forward pass against
the network's closed form in the all-active (linear) regime."""
import numpy as np


def make(F, seed):
    rng = np.random.default_rng(seed)
    from oracle import oracle as o
    P = o.mlp_num_params(F)
    p = (rng.standard_normal(P) * 0.2).astype(np.float32)
    X = rng.standard_normal((12, F)).astype(np.float32)
    Y = rng.standard_normal(12).astype(np.float32)
    return p, X, Y


def test_gradient_matches_finite_differences(oracle):
    F = 7
    p, X, Y = make(F, 0)
    loss, g = oracle.mlp_gradient(p, X, Y)
    rng = np.random.default_rng(1)
    idx = rng.choice(p.size, 60, replace=False)
    idx = np.concatenate([idx, [p.size - 1, p.size - 2, 0, F * 64]])   # b4, w4, W1, b1
    checked = 0
    for i in idx:
        h = np.float32(1e-2) * max(np.float32(1.0), abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] = p[i] + h
        pm[i] = p[i] - h
        lp, _ = oracle.mlp_gradient(pp, X, Y)
        lm, _ = oracle.mlp_gradient(pm, X, Y)
        fd = (lp - lm) / (float(pp[i]) - float(pm[i]))
        # a ReLU kink inside [p-h, p+h] breaks the difference quotient: skip those
        lpp, _ = oracle.mlp_gradient(np.where(np.arange(p.size) == i, p + h / 2, p).astype(np.float32), X, Y)
        lmm, _ = oracle.mlp_gradient(np.where(np.arange(p.size) == i, p - h / 2, p).astype(np.float32), X, Y)
        fd2 = (lpp - lmm) / (float(np.float32(p[i] + h / 2)) - float(np.float32(p[i] - h / 2)))
        if abs(fd - fd2) > 1e-6 * (1 + abs(fd)) + 1e-9:
            continue
        assert abs(fd - g[i]) <= 1e-5 * (1 + abs(g[i])), (i, fd, g[i])
        checked += 1
    assert checked >= 40


def test_forward_linear_regime_closed_form(oracle):
    """All pre-activations positive: y = ((x W1 + b1) W2 + b2) W3 + b3) w4 + b4."""
    F = 5
    rng = np.random.default_rng(3)
    P = oracle.mlp_num_params(F)
    p = np.abs(rng.standard_normal(P)).astype(np.float32) * 0.1   # non-negative weights and biases
    X = np.abs(rng.standard_normal((9, F))).astype(np.float32)
    o = 0
    W1 = p[o:o + F * 64].reshape(F, 64).astype(np.float64); o += F * 64
    b1 = p[o:o + 64]; o += 64
    W2 = p[o:o + 4096].reshape(64, 64).astype(np.float64); o += 4096
    b2 = p[o:o + 64]; o += 64
    W3 = p[o:o + 4096].reshape(64, 64).astype(np.float64); o += 4096
    b3 = p[o:o + 64]; o += 64
    w4 = p[o:o + 64]; o += 64
    b4 = p[o]
    ref = (((X @ W1 + b1) @ W2 + b2) @ W3 + b3) @ w4 + b4
    np.testing.assert_allclose(oracle.mlp_forward(p, X), ref, rtol=1e-12)


def test_gradient_of_zero_residual_is_zero(oracle):
    p, X, _ = make(4, 5)
    Y = oracle.mlp_forward(p, X).astype(np.float32)
    loss, g = oracle.mlp_gradient(p, X, Y)
    assert loss < 1e-12 and np.abs(g).max() < 1e-5
