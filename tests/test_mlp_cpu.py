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
    # f3 / f4: transposes; f5..f7: int A, int B, f64; f8..f12: the family (BF16X3 first);
    # f17: wave-tile aspect log2(WM / WN) / 2 (variant 0: 128 x 64)
    assert F[0, 3] == 0 and F[0, 4] == 1 and F[0, 5] == 0 and F[0, 7] == 0
    assert np.all(F[:, 8] == 1) and np.all(F[:, 9:13] == 0)
    assert F[0, 17] == 0.5 and F[1, 17] == -0.5 and F[2, 17] == 0


@pytest.mark.parametrize("family,count", [(tuner.GEMM_FMA, 4), (tuner.GEMM_F64, 3), (tuner.GEMM_GEMV, 9),
                                          (tuner.GEMM_BF16X6, 5)])
def test_features_of_every_family(family, count):
    s = (3000, 700, 5000, 1, 0, 1)
    F = np.array([tuner.features(s, v, family) for v in range(count)])
    assert np.all(np.isfinite(F)) and np.all(np.abs(F) <= 4)
    assert len({tuple(r) for r in F}) == count
    slot = tuner.FAMILIES.index(family)
    assert np.all(F[:, 8 + slot] == 1) and F[:, 8:13].sum() == count
    assert np.all(F[:, 7] == (family == tuner.GEMM_F64))
    if family == tuner.GEMM_GEMV:
        assert np.all(F[:, 1] == 0)                  # N = 1
    # the families' inputs differ for one shape
    G = tuner.features(s, 0, tuner.GEMM_BF16X3)
    assert not np.array_equal(F[0], G)


def test_features_reject_bad_variant():
    import chai_amd
    with pytest.raises(chai_amd.ChaiError):
        tuner.features((64, 64, 64, 0, 0, 4), 99)
    for family, nv in [(tuner.GEMM_FMA, 4), (tuner.GEMM_F64, 3), (tuner.GEMM_GEMV, 9), (tuner.GEMM_BF16X6, 5)]:
        assert tuner.num_variants(family) == nv
        with pytest.raises(chai_amd.ChaiError):
            tuner.features((64, 64, 64, 0, 0, 4), nv, family)


def test_shape_sampler_range_and_step():
    S = tuner.random_shapes(500, seed=1)
    a = np.array([s[:3] for s in S])
    assert a.min() >= 64 and a.max() <= 16384 and np.all(a % 16 == 0)
    # log-uniform: about half of the sizes below the geometric midpoint 1024
    assert 0.4 < (a < 1024).mean() < 0.6
