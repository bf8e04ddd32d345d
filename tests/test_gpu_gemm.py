"""GPU parity: chai_hip_gemm (FP32-FMA, split-BF16x3 MFMA, FP64 MFMA, mixed
int x float) against the oracle's FP64 GEMM and interpreter restatement.

Tolerances (stated per element, deterministic error bounds):
  FMA     |C - C64| <= gamma_K * sum_k |a_k b_k|,  gamma_K = K u / (1 - K u), u = 2^-24
  BF16X3  |C - C64| <= (3 * 2^-16 + gamma_3K) * sum_k |a_k b_k|   (split residual
          |x - hi - lo| <= 2^-16 |x| per operand: bf16 keeps 8 significant bits,
          so |x - hi| <= 2^-8 |x| and |(x - hi) - lo| <= 2^-8 |x - hi|; the dropped
          lo*lo term is below 2^-16 |a b|; three
          products accumulated per k)
  F64     |C - C64| <= 2 K 2^-53 * sum_k |a_k b_k|
and the interpreter's own result (product rounded to f32, double sum) lies
within u * sum|ab| + u |C| of C64, so GPU-vs-interpreter uses the sum of both.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

U = 2.0 ** -24


def unif(oracle, seed, rows, cols, dtype=np.float32):
    prec = 8 if dtype == np.float64 else 4
    return oracle.rng_fill(1230, prec, 0, seed, cols, rows, 1, limits=(-1.0, 1.0)).reshape(rows, cols)


def bound(mode, K):
    g = K * U / (1 - K * U)
    if mode == "fma":
        return g
    if mode == "bf16x3":
        # three products per k accumulate into one FP32 sum: 3K roundings
        g3 = 3 * K * U / (1 - 3 * K * U)
        return 3 * 2.0 ** -16 + g3
    return 2 * K * 2.0 ** -53


def check(mode, C, A, B, oracle, ta=False, tb=False):
    A64 = (A.T if ta else A).astype(np.float64)
    B64 = (B.T if tb else B).astype(np.float64)
    C64 = A64 @ B64
    absprod = np.abs(A64) @ np.abs(B64)
    K = A64.shape[1]
    err = np.abs(C.astype(np.float64) - C64)
    assert np.all(err <= bound(mode, K) * absprod + 1e-300), (mode, float(err.max()))
    return err


MODES = {"fma": 1, "bf16x3": 2, "f64": 3}


@pytest.mark.parametrize("mode", ["fma", "bf16x3"])
@pytest.mark.parametrize("M,N,K", [(256, 256, 256), (1, 1, 1), (17, 33, 65), (300, 129, 1000),
                                   (512, 768, 96), (1024, 1024, 1024), (512, 512, 1)])
def test_gemm_f32(chai, oracle, mode, M, N, K):
    print("case", mode, M, N, K, flush=True)
    A = unif(oracle, 42, M, K); B = unif(oracle, 43, K, N)
    C = chai.gemm(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), mode=MODES[mode])
    C = C.cpu().numpy()
    check(mode, C, A, B, oracle)
    if M * N * K <= 64 ** 3:
        I = oracle.matmul_interp(A, B)
        tol = bound(mode, K) * (np.abs(A.astype(np.float64)) @ np.abs(B.astype(np.float64)))
        tol = tol + U * (np.abs(A.astype(np.float64)) @ np.abs(B.astype(np.float64))) + U * np.abs(I)
        assert np.all(np.abs(C.astype(np.float64) - I) <= tol)


@pytest.mark.parametrize("mode", ["fma", "bf16x3"])
@pytest.mark.parametrize("ta,tb", [(True, False), (False, True), (True, True)])
def test_gemm_transposes(chai, oracle, mode, ta, tb):
    M, N, K = 130, 70, 300
    A = unif(oracle, 1, K, M) if ta else unif(oracle, 1, M, K)
    B = unif(oracle, 2, N, K) if tb else unif(oracle, 2, K, N)
    C = chai.gemm(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), trans_a=ta, trans_b=tb,
                  mode=MODES[mode]).cpu().numpy()
    check(mode, C, A, B, oracle, ta, tb)


@pytest.mark.parametrize("mode", ["fma", "bf16x3"])
def test_gemm_general_and_batch(chai, oracle, mode):
    M, N, K, nb = 96, 200, 64, 3
    A = np.stack([unif(oracle, 10 + i, M, K) for i in range(nb)])
    B = unif(oracle, 20, K, N)                        # shared B (sameDataB)
    C0 = np.stack([unif(oracle, 30 + i, M, N) for i in range(nb)])
    Ct = torch.from_numpy(C0.copy()).cuda()
    chai.gemm(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), C=Ct, general=True,
              alpha=0.5, beta=-1.5, mode=MODES[mode])
    C = Ct.cpu().numpy()
    for i in range(nb):
        ab = A[i].astype(np.float64) @ B.astype(np.float64)
        ref = 0.5 * ab - 1.5 * C0[i].astype(np.float64)
        absprod = np.abs(A[i].astype(np.float64)) @ np.abs(B.astype(np.float64))
        tol = 0.5 * bound(mode, K) * absprod + U * (np.abs(ref) + 1.5 * np.abs(C0[i])) + 1e-30
        assert np.all(np.abs(C[i] - ref) <= tol)


def test_gemm_sentinel_ones(chai):
    """Autotune sentinel (EvergreenMatmulMM.cpp:800-822): A = B = 1 => C = K exactly."""
    for mode in (1, 2):
        K = 777
        A = torch.ones(300, K, device="cuda"); B = torch.ones(K, 260, device="cuda")
        C = chai.gemm(A, B, mode=mode)
        assert torch.all(C == K).item()


def test_gemm_mixed_int_float(chai, oracle):
    rng = np.random.default_rng(7)
    M = N = K = 512
    Ai = rng.integers(-128, 128, size=(M, K)).astype(np.int32)
    B = unif(oracle, 3, K, N)
    C = chai.gemm(torch.from_numpy(Ai).cuda(), torch.from_numpy(B).cuda(), mode=2,
                  out_dtype=torch.float32).cpu().numpy()
    C64 = Ai.astype(np.float64) @ B.astype(np.float64)
    absprod = np.abs(Ai.astype(np.float64)) @ np.abs(B.astype(np.float64))
    # the int operand is exact in bf16 (|a| <= 256, so its lo panel is zero and
    # the al*bh products are skipped): only B's split residual (2^-16) and the
    # f32 accumulation of 2K products remain
    g2 = 2 * K * U / (1 - 2 * K * U)
    assert np.all(np.abs(C - C64) <= (2.0 ** -16 + g2) * absprod)


def test_gemm_int_not_exact_in_bf16(chai, oracle):
    """i32 values with more than 8 significant bits: lo is nonzero, all three
    products run, the general split bound holds."""
    rng = np.random.default_rng(8)
    M, N, K = 300, 200, 400
    Ai = rng.integers(-(1 << 20), 1 << 20, size=(M, K)).astype(np.int32)
    B = unif(oracle, 4, K, N)
    C = chai.gemm(torch.from_numpy(Ai).cuda(), torch.from_numpy(B).cuda(), mode=2,
                  out_dtype=torch.float32).cpu().numpy()
    A64 = Ai.astype(np.float64)
    C64 = A64 @ B.astype(np.float64)
    absprod = np.abs(A64) @ np.abs(B.astype(np.float64))
    assert np.all(np.abs(C - C64) <= bound("bf16x3", K) * absprod)


@pytest.mark.parametrize("M,N,K", [(256, 256, 1024), (300, 200, 77), (1000, 1100, 600)])
def test_gemm_bf16_exact_operands_are_exact(chai, M, N, K):
    """Integer-valued f32 operands in [-8, 8] are exact in bf16: both lo panels
    are zero, every tile runs the hi*hi product alone, and with |C| < 2^24 every
    partial sum is exact, so C equals the FP64 product bit for bit."""
    rng = np.random.default_rng(M + N + K)
    A = rng.integers(-8, 9, size=(M, K)).astype(np.float32)
    B = rng.integers(-8, 9, size=(K, N)).astype(np.float32)
    C = chai.gemm(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), mode=2).cpu().numpy()
    C64 = A.astype(np.float64) @ B.astype(np.float64)
    assert np.array_equal(C.astype(np.float64), C64)


def test_gemm_zero_lo_panels_per_tile(chai, oracle):
    """Tiles decide separately: A is bf16-exact except in rows 300..309 (one
    tile row), B is general.  Tiles of exact A rows skip the al*bh products;
    every tile meets the split bound against FP64."""
    rng = np.random.default_rng(5)
    M, N, K = 1024, 512, 700
    A = rng.integers(-100, 101, size=(M, K)).astype(np.float32)
    A[300:310] = unif(oracle, 9, 10, K)
    B = unif(oracle, 10, K, N)
    C = chai.gemm(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), mode=2).cpu().numpy()
    A64, B64 = A.astype(np.float64), B.astype(np.float64)
    err = np.abs(C - A64 @ B64)
    assert np.all(err <= bound("bf16x3", K) * (np.abs(A64) @ np.abs(B64)))


@pytest.mark.parametrize("M,N,K", [(128, 128, 128), (65, 97, 33), (512, 384, 640)])
def test_dgemm(chai, oracle, M, N, K):
    A = unif(oracle, 5, M, K, np.float64); B = unif(oracle, 6, K, N, np.float64)
    C = chai.gemm(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), mode=3).cpu().numpy()
    check("f64", C, A, B, oracle)
    ref = oracle.gemm_f64ref(A, B, nthreads=4)
    assert np.all(np.abs(C - ref) <= 4 * K * 2.0 ** -53 * (np.abs(A) @ np.abs(B)))


def test_dgemm_variants_agree(chai, oracle):
    """Every FP64 tile shape accumulates each output over k in the same order
    (4-deep MFMA steps), so all variants give the same bits; the default choice
    is one of them."""
    M, N, K = 300, 520, 200
    A = unif(oracle, 11, M, K, np.float64); B = unif(oracle, 12, K, N, np.float64)
    At, Bt = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    outs = [chai.gemm(At, Bt, mode=MODES["f64"], variant=v) for v in range(len(chai.gemm_variants(MODES["f64"])))]
    assert len(outs) == 3
    for o in outs[1:]:
        assert torch.equal(o, outs[0])
    assert torch.equal(chai.gemm(At, Bt, mode=MODES["f64"]), outs[0])
    check("f64", outs[0].cpu().numpy(), A, B, oracle)
    # transposed operands (packed k-major on the way in) give the same bits
    Ct = chai.gemm(At.t().contiguous(), Bt.t().contiguous(), trans_a=True, trans_b=True, mode=MODES["f64"], variant=1)
    assert torch.equal(Ct, outs[0])


def test_gemm_8192_sampled_rows(chai, oracle):
    """The headline shape, bf16x3 and FMA, checked on 8 sampled rows against FP64."""
    n = 8192
    A = unif(oracle, 42, n, n); B = unif(oracle, 43, n, n)
    At, Bt = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    rows = np.random.default_rng(0).choice(n, 8, replace=False)
    Bd = B.astype(np.float64)
    C64 = A[rows].astype(np.float64) @ Bd
    absprod = np.abs(A[rows].astype(np.float64)) @ np.abs(Bd)
    for name, mode in (("fma", 1), ("bf16x3", 2)):
        C = chai.gemm(At, Bt, mode=mode)[torch.from_numpy(rows).cuda()].cpu().numpy()
        err = np.abs(C - C64)
        assert np.all(err <= bound(name, n) * absprod), (name, float(err.max()))
        print(name, "max abs err vs FP64", float(err.max()), "rms", float(np.sqrt((err ** 2).mean())))


def test_gemm_rejects_bad_arguments(chai):
    A = torch.ones(4, 4, device="cuda")
    with pytest.raises(ValueError):
        chai.gemm(A, torch.ones(5, 4, device="cuda"))  # inner dimension mismatch (binding check)
    Ai = torch.ones(4, 4, dtype=torch.int32, device="cuda")
    with pytest.raises(chai.ChaiError):
        chai.gemm(Ai, Ai, out_dtype=torch.float32)


def _bf16_rne(x):
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32)


def test_bf16x3_rank1_matches_split_model(chai, oracle):
    """K = 1: C = (ah*bl + al*bh) + ah*bh with hi = bf16_rne(x), lo = bf16_rne(x - hi),
    each product exact in FP32.  The GPU must agree with this model to one
    FP32 rounding per accumulation (it pins the hardware conversion as round to
    nearest even and the three-product order)."""
    M, N = 256, 256
    A = unif(oracle, 7, M, 1)
    B = unif(oracle, 8, 1, N)
    C = chai.gemm(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), mode=MODES["bf16x3"]).cpu().numpy()
    ah = _bf16_rne(A); al = _bf16_rne(A - ah)
    bh = _bf16_rne(B); bl = _bf16_rne(B - bh)
    p1 = (ah @ bl).astype(np.float32)          # rank-1 products, exact in FP32
    p2 = (al @ bh).astype(np.float32)
    p3 = (ah @ bh).astype(np.float32)
    model = (p1 + p2).astype(np.float32) + p3
    ulp = np.spacing(np.abs(model).astype(np.float32))
    dev = np.abs(C.astype(np.float64) - model.astype(np.float64)) / ulp
    print("max deviation from the split model: %.2f ulp" % dev.max())
    assert dev.max() <= 1.0


def test_tuner_journal_crash_resume(chai, tmp_path):
    """Crash-resume through the journal (autotune/AutoTuneMemo.cpp:101-188):
    a tuning run journals every variant; a run that died while compiling the
    winner leaves PRINT_SOURCE, COMPILE_KERNEL as that kernel's last lines.
    Reloading such a journal blacklists the winner, reuses every other
    variant's timing without running it, and the search picks the best of
    the rest."""
    import torch
    n = 512
    A, B = torch.rand(n, n, device="cuda"), torch.rand(n, n, device="cuda")
    C = torch.empty(n, n, device="cuda")
    first = tmp_path / "journal0.AMD"
    chai.memo_close()
    chai.memo_open(first, 0)
    v1, _ = chai.gemm_tune(A, B, C, trials=1)
    chai.memo_close()
    lines = [l.split() for l in first.read_text().split("\n") if l.strip()]
    nv = len(chai.gemm_variants())
    by_v = {v: [l for l in lines if l[1].endswith("_v%d" % v)] for v in range(nv)}
    assert all(by_v[v] for v in range(nv))
    times = {v: next(int(r[2]) for r in by_v[v] if int(r[2]) > 0) for v in range(nv)}
    assert all(t > 0 for t in times.values())
    # output checks (BEFORE_CHECK -100, AFTER_CHECK -101): the first variant run
    # passes the FP64 check of sampled rows and becomes the reference, every
    # other one is compared with it, and none differs
    assert all([r[2] for r in by_v[v]][-2:] == ["-100", "-101"] for v in range(nv))
    assert not any(r[2] == "-7" for r in lines)
    crashed = tmp_path / "journal1.AMD"
    with open(crashed, "w") as f:
        for v in range(nv):
            rows = by_v[v] if v != v1 else [r for r in by_v[v] if r[2] in ("-1", "-2")]
            for r in rows:
                f.write(" ".join(r) + "\n")
    before = crashed.read_text()
    chai.memo_open(crashed, 0)
    try:
        name = by_v[v1][0][1]
        assert chai.memo_status(name)[0] == -1
        v2, ms2 = chai.gemm_tune(A, B, C, trials=1)
        rest = {v: t for v, t in times.items() if v != v1}
        assert v2 == min(rest, key=rest.get) and abs(ms2 * 1000.0 - rest[v2]) <= 1.0
        assert crashed.read_text() == before          # memo hits only: nothing re-timed
    finally:
        chai.memo_close()


@pytest.mark.parametrize("mode", ["bf16x3", "f64"])
def test_tuner_checks_every_variant_against_the_first(chai, tmp_path, mode):
    """The tuning search compares each variant's output with the first one's
    (every variant accumulates in the same k order; AutoTuneBench's
    checkOutput).  Integer-valued A (all lo panels zero) makes some tiles skip
    products, which may only change the sign of zeros: compared by value."""
    import torch
    M, N, K = 300, 520, 200
    dt = torch.float64 if mode == "f64" else torch.float32
    A = torch.randint(-8, 8, (M, K), device="cuda").to(dt)
    A[:100] = torch.rand(100, K, device="cuda", dtype=dt)        # lo panels nonzero in the first rows only
    B = torch.rand(K, N, device="cuda", dtype=dt)
    C = torch.empty(M, N, device="cuda", dtype=dt)
    j = tmp_path / "journal0.AMD"
    chai.memo_close()
    chai.memo_open(j, 0)
    try:
        m = chai.GEMM_F64 if mode == "f64" else chai.GEMM_BF16X3
        chai.gemm_tune(A, B, C, mode=m, trials=1)
    finally:
        chai.memo_close()
    lines = [l.split() for l in j.read_text().split("\n") if l.strip()]
    states = [r[2] for r in lines]
    assert "-7" not in states
    assert states.count("-101") == len(chai.gemm_variants(m))


def test_tuner_skips_a_variant_journalled_bad_output(chai, tmp_path):
    """A variant whose output check failed in an earlier run (BEFORE_CHECK ->
    BAD_OUTPUT in the journal, autotune/AutoTuneMemo.cpp:101-188) stays
    blacklisted after a reload: the search neither runs nor picks it."""
    import torch
    n = 384
    A, B = torch.rand(n, n, device="cuda"), torch.rand(n, n, device="cuda")
    C = torch.empty(n, n, device="cuda")
    first = tmp_path / "journal0.AMD"
    chai.memo_close()
    chai.memo_open(first, 0)
    try:
        v1, _ = chai.gemm_tune(A, B, C, trials=1)
    finally:
        chai.memo_close()
    lines = [l.split() for l in first.read_text().split("\n") if l.strip()]
    name = next(r[1] for r in lines if r[1].endswith("_v%d" % v1))
    # the winner's last check failed: its times are dropped, BAD_OUTPUT follows BEFORE_CHECK
    j = tmp_path / "journal1.AMD"
    with open(j, "w") as f:
        for r in lines:
            if r[1] == name and int(r[2]) > 0:
                continue
            if r[1] == name and r[2] == "-101":
                f.write("%s %s -7\n" % (r[0], r[1]))
                continue
            f.write(" ".join(r) + "\n")
    chai.memo_open(j, 0)
    try:
        assert chai.memo_status(name)[0] == -1
        before = j.read_text()
        v2, _ = chai.gemm_tune(A, B, C, trials=1)
        assert v2 != v1
        assert j.read_text() == before            # nothing re-run, the bad variant included
    finally:
        chai.memo_close()
    chai.memo_open(j, 0)                          # and it is still bad after another reload
    try:
        assert chai.memo_status(name)[0] == -1
    finally:
        chai.memo_close()
