"""GPU: the MI355X device inside the reference's own runtime.

The reference-API programs of tests/frontend/ and the reference's sample
programs (built here from its sources) run twice on the GPU machine: with the
reference's interpreter only, and with the MI355X device over
libchai_mi355x.so (CHAI_MI355X=0), which takes every trace it can lower.
Results are compared with the interpreter's:

- integer arrays, IEEE-exact elementwise ops (+ - * / sqrt fma floor, compares),
  gathers, index leaves, transposes: bit-exact;
- GEMM in the split-BF16x3 form: |C - C_interp| <= (3 2^-16 + gamma_3K) sum|a b|
  (the interpreter accumulates in double and rounds once, so its value is
  within 2^-24 |C| of exact); FP64 GEMM: 2K 2^-53 sum|a b|;
- reductions: the FP64 partial sums combine in another order than the
  interpreter's sequential double sum, so the f32 result may differ by one
  rounding: |r - r_interp| <= 2^-23 |r_interp| + n 2^-52 sum|x|.
Integer x integer products run on the GPU's exact integer GEMM: bit-exact.
"""
import re

import numpy as np
import pytest

from tests import fe_programs as fe

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not fe.built(), reason="reference-API programs not built (needs the reference tree)")]


def _inputs(case):
    return fe.case_inputs(case)


def _gemm_bound(A, B, K, split=True):
    mag = np.abs(A.astype(np.float64)) @ np.abs(B.astype(np.float64))
    u = 2.0 ** -24
    coef = (3 * 2.0 ** -16 + (3 * K + 2) * u) if split else (2 * K + 2) * 2.0 ** -53
    return coef * mag + 2 * u * mag


def _run(case, tmp_path):
    print("frontend case %s" % case, flush=True)
    ref, _ = fe.run_case(case, str(tmp_path / "interp"))
    got, err = fe.run_case(case, str(tmp_path / "gpu"), device="gpu")
    assert "not available" not in err, err
    return ref, got, err


def test_mm_f32_config1(tmp_path):
    """config 1: C = matmul(A, B), 256^3, inputs from Chai's Philox stream (seed 42)"""
    ref, got, err = _run("mm_f32", tmp_path)
    assert "refused" not in err, err
    x = _inputs("mm_f32")
    A, B = x["A"].reshape(256, 256), x["B"].reshape(256, 256)
    bound = _gemm_bound(A, B, 256)
    d = np.abs(got["C"].astype(np.float64) - ref["C"].astype(np.float64)).reshape(256, 256)
    assert np.all(d <= bound), d.max()


def test_mm_shapes(tmp_path):
    ref, got, err = _run("mm_shapes", tmp_path)
    assert "refused" not in err, err
    x = _inputs("mm_shapes")
    M, K, N = 96, 160, 72
    A, At, B, Bt = x["A"].reshape(M, K), x["At"].reshape(K, M), x["B"].reshape(K, N), x["Bt"].reshape(N, K)
    C0 = x["C0"].reshape(M, N)
    for name, a, b in [("C1", At.T, B), ("C2", A, Bt.T), ("C3", At.T, Bt.T)]:
        d = np.abs(got[name].astype(np.float64) - ref[name].astype(np.float64)).reshape(M, N)
        assert np.all(d <= _gemm_bound(a, b, K)), (name, d.max())
    # general form: 0.75 A B - 1.5 C0
    d = np.abs(got["C4"].astype(np.float64) - ref["C4"].astype(np.float64)).reshape(M, N)
    assert np.all(d <= 0.75 * _gemm_bound(A, B, K) + 2.0 ** -23 * (1.5 * np.abs(C0) + 1)), d.max()
    # matrix-vector: FP32 lanes, FP64 combine
    y = x["x"].astype(np.float64)
    mag = np.abs(A.astype(np.float64)) @ np.abs(y)
    d = np.abs(got["y"].astype(np.float64) - ref["y"].astype(np.float64))
    assert np.all(d <= (K + 2) * 2.0 ** -24 * mag), d.max()


def test_mm_int_exact_and_mixed(tmp_path):
    ref, got, err = _run("mm_int", tmp_path)
    assert "refused" not in err, err
    for name in ("CUU", "CII", "CUI"):
        assert np.array_equal(ref[name], got[name]), name
    x = _inputs("mm_int")
    IA, FB = x["IA"].reshape(64, 48), x["FB"].reshape(48, 40)
    d = np.abs(got["CIF"].astype(np.float64) - ref["CIF"].astype(np.float64)).reshape(64, 40)
    assert np.all(d <= _gemm_bound(IA.astype(np.float32), FB, 48)), d.max()


def test_mm_f64(tmp_path):
    ref, got, err = _run("mm_f64", tmp_path)
    assert "refused" not in err, err
    x = _inputs("mm_f64")
    A, B = x["A"].reshape(64, 80), x["B"].reshape(80, 48)
    d = np.abs(got["C"] - ref["C"]).reshape(64, 48)
    assert np.all(d <= _gemm_bound(A, B, 80, split=False)), d.max()


@pytest.mark.parametrize("case", ["ew", "ints", "gather"])
def test_elementwise_bit_exact(case, tmp_path):
    ref, got, err = _run(case, tmp_path)
    assert "refused" not in err, err
    for name in ref:
        assert np.array_equal(ref[name].view(np.uint8), got[name].view(np.uint8)), (case, name)


def _red_close(got, ref, n, mag):
    return abs(float(got) - float(ref)) <= 2.0 ** -23 * abs(float(ref)) + n * 2.0 ** -52 * mag


def test_reductions(tmp_path):
    ref, got, err = _run("red", tmp_path)
    assert "refused" not in err, err
    x = _inputs("red")
    A, B = x["A"].astype(np.float64), x["B"].astype(np.float64)
    n = A.size
    mags = [np.abs(A).sum(), 2 * np.abs(B).sum() / n, np.abs(A * B).sum(), (A * A + np.abs(B)).sum()]
    for i in range(4):
        assert _red_close(got["r"][i], ref["r"][i], n, mags[i]), (i, got["r"][i], ref["r"][i])
    assert np.array_equal(got["ri"], ref["ri"])


def test_multi_trace(tmp_path):
    ref, got, err = _run("multi", tmp_path)
    assert "refused" not in err, err
    assert np.array_equal(ref["D"], got["D"]) and np.array_equal(ref["E"], got["E"])
    E, D = ref["E"].astype(np.float64), ref["D"].astype(np.float64)
    assert _red_close(got["s"][0], ref["s"][0], D.size, np.abs(D).sum())
    assert _red_close(got["s"][1], ref["s"][1], E.size, np.abs(E - D).sum())


_NUM = re.compile(r"-?\d+\.?\d*(?:e[-+]?\d+)?")


@pytest.mark.parametrize("sample", fe.SAMPLES)
def test_reference_samples(sample):
    """The reference's sample programs, unchanged: the printed results (6
    significant digits) of the GPU run equal the interpreter's, up to one unit
    in the last printed digit for floating-point sums and products.  The
    random samples (pi, monte) are compared with the same lowering over the
    CPU oracle instead: the interpreter's generator is not Chai's Random123
    stream and changes from run to run."""
    print("reference sample %s" % sample, flush=True)
    ref, _ = fe.run_sample(sample, device="oracle" if sample in fe.RNG_SAMPLES else None)
    got, err = fe.run_sample(sample, device="gpu")
    assert "not available" not in err and "refused" not in err, err
    if sample == "md5":
        assert got == ref
        return
    rl, gl = ref.strip().splitlines(), got.strip().splitlines()
    assert len(rl) == len(gl) and rl
    for a, b in zip(rl, gl):
        na, nb = [float(v) for v in _NUM.findall(a)], [float(v) for v in _NUM.findall(b)]
        assert len(na) == len(nb)
        for u, v in zip(na, nb):
            assert abs(u - v) <= 1e-5 * max(abs(u), 1.0), (a, b)
