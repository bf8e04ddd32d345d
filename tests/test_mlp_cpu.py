"""Host-side parts of the learned tuner that need no GPU: the GEMM feature map
(chai_hip_gemm_features) and the shape sampler."""
import numpy as np
import pytest

from chai_amd import tuner


def test_features_are_scaled_and_distinguish_variants():
    s = (4096, 1024, 777, 0, 1, 4)
    F = np.array([tuner.features(s, v) for v in range(8)])
    assert F.shape == (8, tuner.NFEAT)
    assert np.all(np.isfinite(F)) and np.all(np.abs(F) <= 4)
    assert len({tuple(r) for r in F}) == 8           # every variant gets its own input
    np.testing.assert_allclose(F[:, :3], [[12 / 14, 10 / 14, np.log2(777) / 14]] * 8, rtol=1e-6)
    # f3: any transposed operand; f4: wave-tile aspect log2(WM / WN) / 2 (variant 0: 128 x 64)
    assert F[0, 3] == 1 and F[0, 4] == 0.5 and F[1, 4] == -0.5 and F[2, 4] == 0 and F[0, 5] == 0


def test_features_reject_bad_variant():
    import chai_amd
    with pytest.raises(chai_amd.ChaiError):
        tuner.features((64, 64, 64, 0, 0, 4), 99)


def test_shape_sampler_range_and_step():
    S = tuner.random_shapes(500, seed=1)
    a = np.array([s[:3] for s in S])
    assert a.min() >= 64 and a.max() <= 16384 and np.all(a % 16 == 0)
    # log-uniform: about half of the sizes below the geometric midpoint 1024
    assert 0.4 < (a < 1024).mean() < 0.6
