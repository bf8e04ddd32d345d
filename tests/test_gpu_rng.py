"""GPU parity: chai_hip_rng_fill against the oracle's stream, bit-exact.

Integer words, uniforms and normals are all compared bit for bit: the
normals' log / sin / cos are the stream's own routines on both sides, in
double for the f64 streams and in f32 for the f32 streams (see oracle/rng.c)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

TORCH_DT = {0: torch.int32, 1: torch.int32, 4: torch.float32, 8: torch.float64}


def gpu_fill(chai, variant, prec, vl, seed, W, H, S, normal=False, limits=None):
    out = torch.empty(S * H * W, dtype=TORCH_DT[prec], device="cuda")
    chai.rng_fill(out, variant=variant, seed=seed, vec_len=vl, normal=normal, limits=limits,
                  W=W, H=H, slots=S, prec=prec)
    torch.cuda.synchronize()
    a = out.cpu().numpy().reshape(S, H, W)
    return a.view(np.uint32) if prec == 0 else a


CASES = [(v, p, vl) for v in (1230, 1231) for p, vls in ((0, (1, 2, 4)), (1, (4,)), (4, (1, 2, 4)),
                                                          (8, (1, 2))) for vl in vls]


@pytest.mark.parametrize("variant,prec,vl", CASES)
@pytest.mark.parametrize("shape", [(7, 3, 2), (64, 16, 1), (1, 1, 1), (1000, 3, 3)])
def test_uniform_bit_exact(chai, oracle, variant, prec, vl, shape):
    W, H, S = shape
    print("case", variant, prec, vl, shape, flush=True)
    g = gpu_fill(chai, variant, prec, vl, 12345, W, H, S)
    o = oracle.rng_fill(variant, prec, vl, 12345, W, H, S)
    assert np.array_equal(g.view(np.uint8), o.view(np.uint8))


@pytest.mark.parametrize("variant", [1230, 1231])
@pytest.mark.parametrize("prec,vl", [(4, 4), (4, 2), (4, 1), (8, 2), (8, 1)])
def test_normal_bit_exact(chai, oracle, variant, prec, vl):
    for (W, H, S) in [(13, 5, 2), (256, 64, 1)]:
        g = gpu_fill(chai, variant, prec, vl, 7, W, H, S, normal=True)
        o = oracle.rng_fill(variant, prec, vl, 7, W, H, S, normal=True)
        assert np.array_equal(g.view(np.uint8), o.view(np.uint8))


@pytest.mark.parametrize("prec", [4, 8])
def test_limits_bit_exact(chai, oracle, prec):
    for lim in [(-1.0, 1.0), (0.0, 3.5), (2.0, 5.0)]:
        g = gpu_fill(chai, 1230, prec, 0, 42, 100, 10, 1, limits=lim)
        o = oracle.rng_fill(1230, prec, 0, 42, 100, 10, 1, limits=lim)
        assert np.array_equal(g, o)


def test_full_size_normals_2_24(chai, oracle):
    """The Monte Carlo workload's 2^24 Philox4x32-10 normals (seed 7), bit-exact."""
    n = 1 << 24
    g = gpu_fill(chai, 1230, 4, 4, 7, n, 1, 1, normal=True)
    o = oracle.rng_fill(1230, 4, 4, 7, n, 1, 1, normal=True)
    assert np.array_equal(g.view(np.uint32), o.view(np.uint32))
    x = g.astype(np.float64).ravel()
    assert abs(x.mean()) < 5.0 / np.sqrt(n) and abs(x.var() - 1.0) < 2e-3


def test_normals_short_forms_exhaustive(chai):
    """The f32 normals' log (branch-free, reciprocal-plus-correction divide)
    and sqrt (hardware root moved by fma residuals) give the bits of bm_logf
    with its IEEE divide and of the correctly rounded sqrtf for every one of
    the 2^32 stream words -- the whole domain unit_f32 can feed them."""
    print("case exhaustive normals check", flush=True)
    assert chai.rng_check_normals() == (0, 0)


def test_rejects_bad_arguments(chai):
    out = torch.empty(16, dtype=torch.int32, device="cuda")
    with pytest.raises(chai.ChaiError):
        chai.rng_fill(out, vec_len=3, prec=0, W=16, H=1)
    with pytest.raises(chai.ChaiError):
        chai.rng_fill(out, normal=True, prec=1, W=16, H=1)
    with pytest.raises(chai.ChaiError):
        chai.rng_fill(out, variant=99, prec=0, W=16, H=1)
