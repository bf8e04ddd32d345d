"""The measured workloads, defined once for the benchmark and the tests.

  mc_trace     fused Monte Carlo trace: 2^24 Philox4x32-10 normals (seed 7)
               driving S = 100 exp((0.05 - 0.02) * 1 + 0.2 z), payoff
               max(S - 105, 0), mean with an FP64 cross-block combine
  ragged_sizes 4096 square GEMM edges, log-uniform in [32, 1024], multiples
               of 16 (seed 42), for the ragged batch C = A B + max(C*0.5+1, 0)
  filter_trace the reference's sample/filter.cpp: a 5-point stencil of
               gather2_floor reads, periodic through the subscript wrap
"""
import numpy as np

from . import F32
from .trace import Trace, exp, f32, gather2_floor, maxe, mean


def mc_trace(W, H=1, seed=7):
    """Float literals keep the arithmetic in f32 (f32 scalars, chai.h's float overloads)."""
    t = Trace()
    z = t.rng_normal(F32, W, H, seed=seed)
    drift = float(np.float32(np.float32(0.05) - np.float32(0.02)) * np.float32(1))
    S = f32(100.0) * exp(f32(drift) + f32(0.2) * z)
    m = t.assign(t.output(F32, 1), mean(maxe(S - f32(105.0), f32(0.0))))
    return t, m


def ragged_sizes(count=4096, lo=32, hi=1024, step=16, seed=42):
    rng = np.random.default_rng(seed)
    n = np.exp(rng.uniform(np.log(lo), np.log(hi), count))
    return np.clip((n // step * step).astype(int), lo, hi)


RAGGED_EPILOGUE = (0.5, 1.0, 0.0)   # max(C * 0.5f + 1, 0)


def filter_trace(N):
    """sample/filter.cpp: B = the five gather2_floor reads of A at (X, Y),
    (X - 1, Y), (X + 1, Y), (X, Y - 1), (X, Y + 1), X and Y index_f32 arrays;
    x = -1 wraps to W - 1 and x = W to 0."""
    t = Trace()
    A = t.input(F32, N, N)
    X = Trace.index(F32, 0, N, N)
    Y = Trace.index(F32, 1, N, N)
    B = t.assign(t.output(F32, N, N),
                 gather2_floor(A, X, Y) + gather2_floor(A, X - 1, Y) + gather2_floor(A, X + 1, Y)
                 + gather2_floor(A, X, Y - 1) + gather2_floor(A, X, Y + 1))
    return t, A, B


def filter_expected(a):
    """[..., H, W] -> the same stencil with numpy's periodic rolls, summed in the trace's order"""
    r = lambda k, ax: np.roll(a, k, axis=ax)  # noqa: E731
    return (((a + r(1, -1)) + r(-1, -1)) + r(1, -2)) + r(-1, -2)
