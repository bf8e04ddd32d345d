"""GPU parity: chai_hip_gemm (FP32-FMA, split-BF16x3 MFMA, FP64 MFMA, mixed
int x float) against the oracle's FP64 GEMM and interpreter restatement.

Tolerances (stated per element, deterministic error bounds):
  FMA     |C - C64| <= gamma_K * sum_k |a_k b_k|,  gamma_K = K u / (1 - K u), u = 2^-24
  BF16X3  |C - C64| <= (3 * 2^-16 + gamma_3K) * sum_k |a_k b_k|   (split residual
          |x - hi - lo| <= 2^-16 |x| per operand: bf16 keeps 8 significant bits,
          so |x - hi| <= 2^-8 |x| and |(x - hi) - lo| <= 2^-8 |x - hi|; the dropped
          lo*lo term is below 2^-16 |a b|; three
          products accumulated per k)
  BF16X6  |C - C64| <= (6K + 4) 2^-24 * sum_k |a_k b_k|   (three pieces carry all 24 bits;
          the dropped products of order 3 and 4 are below 2^-24 |a b|; 6K
          accumulations in two passes and one final add)
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
    if mode == "bf16x6":
        return (6 * K + 4) * U
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


MODES = {"fma": 1, "bf16x3": 2, "f64": 3, "bf16x6": 5}


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
    with pytest.raises(chai.ChaiError):     # integer x integer runs on the integer GEMM only
        chai.gemm(Ai, Ai, mode=chai.GEMM_BF16X3, out_dtype=torch.float32)
    with pytest.raises(chai.ChaiError):     # and the integer GEMM takes integer operands only
        chai.gemm(A, A, mode=chai.GEMM_INT)
    # bf16x3 panels hold at most 2^25 padded rows: refused before any allocation
    tall = torch.ones((1 << 25) + 1, 1, device="cuda")
    with pytest.raises(chai.ChaiError):
        chai.gemm(tall, torch.ones(1, 1, device="cuda"), mode=chai.GEMM_BF16X3)
    del tall


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


def _int_operand(rng, shape, kind):
    if kind == "u32":
        return rng.integers(0, 2**32, size=shape, dtype=np.uint64).astype(np.uint32)
    return rng.integers(-2**31, 2**31, size=shape, dtype=np.int64).astype(np.int32)


def _dev(a):
    """device tensor of a numpy u32 / i32 array (u32 travels as int32 bits)"""
    import torch
    return torch.from_numpy(a.view(np.int32)).cuda()


@pytest.mark.parametrize("ka,kb", [("u32", "u32"), ("i32", "i32"), ("u32", "i32"), ("i32", "u32")])
@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True)])
def test_gemm_int_exact_interpreter_semantics(chai, oracle, ka, kb, ta, tb):
    """integer x integer GEMM (CHAI_GEMM_INT): the interpreter's semantics,
    pinned against the reference's own interpreter (tests/test_frontend_cpu.py):
    products wrap in the promoted type, the double sum converts to C the x86
    way (int32 overflow gives INT_MIN).  Bit-exact against the oracle, on
    full-range operands (most int32 sums overflow) and on small ones."""
    import torch
    rng = np.random.default_rng(hash((ka, kb, ta, tb)) & 0xffff)
    M, N, K = 70, 45, 130
    for small in (False, True):
        A = _int_operand(rng, (K, M) if ta else (M, K), ka)
        B = _int_operand(rng, (N, K) if tb else (K, N), kb)
        if small:
            A = (A % 97).astype(A.dtype)
            B = (B % 89).astype(B.dtype)
        ref = oracle.matmul_interp(A, B, trans_a=ta, trans_b=tb)
        pa = chai.U32 if ka == "u32" else chai.I32
        pb = chai.U32 if kb == "u32" else chai.I32
        # the result type is the promoted one: u32 only for u32 x u32 (u32 travels as int32 bits)
        C = chai.gemm(_dev(A), _dev(B), trans_a=ta, trans_b=tb, prec_a=pa, prec_b=pb,
                      out_dtype=torch.int32, prec_c=max(pa, pb))
        got = C.cpu().numpy().view(ref.dtype)
        assert np.array_equal(got, ref), (ka, kb, ta, tb, small)


def test_gemm_int_general_batched_and_long_k(chai, oracle):
    """matmulG on integers (tmp in the promoted type, then alpha tmp + beta C
    in double), a batch sharing A, and K past 2^21 where the kernel sums in
    double in k order as the interpreter does."""
    import torch
    rng = np.random.default_rng(11)
    M, N, K = 33, 40, 50
    A = _int_operand(rng, (M, K), "i32") % 1000
    A = A.astype(np.int32)
    B = _int_operand(rng, (K, N), "i32") % 1000
    B = B.astype(np.int32)
    C0 = _int_operand(rng, (M, N), "i32") % 5000
    C0 = C0.astype(np.int32)
    ref = oracle.matmul_interp(A, B, C=C0, alpha=1.5, beta=-2.0)
    Cd = _dev(C0.copy())
    chai.gemm(_dev(A), _dev(B), C=Cd, general=True, alpha=1.5, beta=-2.0)
    assert np.array_equal(Cd.cpu().numpy(), ref)
    # batch of 3 instances of B against one A
    Bs = np.stack([(_int_operand(rng, (K, N), "u32") % 7919).astype(np.uint32) for _ in range(3)])
    Cb = chai.gemm(_dev(A), _dev(Bs.reshape(-1, N)).reshape(3, K, N), prec_b=chai.U32, out_dtype=torch.int32)
    for i in range(3):
        assert np.array_equal(Cb[i].cpu().numpy(), oracle.matmul_interp(A, Bs[i]))
    # K = 2^21 + 96: the double accumulator rounds; emulated in k order
    K2 = (1 << 21) + 96
    A2 = _int_operand(rng, (2, K2), "u32")
    B2 = _int_operand(rng, (K2, 3), "u32")
    ref2 = oracle.matmul_interp(A2, B2)
    C2 = chai.gemm(_dev(A2), _dev(B2), prec_a=chai.U32, prec_b=chai.U32, out_dtype=torch.int32,
                   prec_c=chai.U32)
    assert np.array_equal(C2.cpu().numpy().view(np.uint32), ref2)


def test_gemv_int_exact(chai, oracle):
    import torch
    rng = np.random.default_rng(12)
    M, N = 300, 257
    A = _int_operand(rng, (M, N), "i32")
    Ad = _dev(A)
    for ta in (False, True):
        xv = (_int_operand(rng, (M if ta else N,), "i32") % 1001).astype(np.int32)
        ref = oracle.matvec_interp(A, xv, trans_a=ta)
        y = chai.gemv(Ad, _dev(xv), trans_a=ta)
        assert y.dtype == torch.int32
        assert np.array_equal(y.cpu().numpy(), ref), ta


def test_config1_against_the_interpreters_result(chai, oracle):
    """config 1 (f32 256^3, A and B from Chai's Philox stream, seed 42): the
    GPU's product against the reference interpreter's own result, kept in
    tests/golden/interp_fixtures.npz.  The interpreter sums in double and
    rounds once (within 2^-24 |C| of exact), so |C - C_interp| <=
    (3 2^-16 + gamma_3K + 2^-23) sum|a b| for the split-BF16x3 family, and
    gamma_(K+1) + 2^-23 for the FP32-FMA one."""
    import os
    import torch
    fix = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "interp_fixtures.npz"))
    ref = fix["mm_f32.out.C"].reshape(256, 256).astype(np.float64)
    A = oracle.rng_fill(1230, 4, 4, 42, 256, 256, limits=(-1.0, 1.0)).reshape(256, 256)
    B = oracle.rng_fill(1230, 4, 4, 42 + 65536, 256, 256, limits=(-1.0, 1.0)).reshape(256, 256)
    mag = np.abs(A.astype(np.float64)) @ np.abs(B.astype(np.float64))
    K = 256
    for mode, coef in [(chai.GEMM_BF16X3, bound("bf16x3", K) + 2.0 ** -23),
                       (chai.GEMM_FMA, bound("fma", K) + 2.0 ** -23)]:
        C = chai.gemm(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), mode=mode).cpu().numpy()
        assert np.all(np.abs(C - ref) <= coef * mag), mode


@pytest.mark.parametrize("M,N,K,ta,tb", [(256, 256, 256, False, False), (300, 520, 200, True, False),
                                         (129, 77, 1000, False, True), (512, 384, 4096, False, False)])
def test_gemm_bf16x6_bound_and_variants(chai, oracle, M, N, K, ta, tb):
    """Split-BF16x6: within (6K + 4) 2^-24 sum|ab| of the FP64 product, and every
    tile variant gives the same C (same k order for each output)."""
    A = unif(oracle, 21, K if ta else M, M if ta else K)
    B = unif(oracle, 22, N if tb else K, K if tb else N)
    At, Bt = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    A64 = (A.T if ta else A).astype(np.float64)
    B64 = (B.T if tb else B).astype(np.float64)
    C64, mag = A64 @ B64, np.abs(A64) @ np.abs(B64)
    first = None
    for v in range(len(chai.gemm_variants(chai.GEMM_BF16X6))):
        C = chai.gemm(At, Bt, trans_a=ta, trans_b=tb, mode=chai.GEMM_BF16X6, variant=v).cpu().numpy()
        assert np.all(np.abs(C - C64) <= bound("bf16x6", K) * mag), v
        if first is None:
            first = C
        else:
            assert np.array_equal(C, first), v


def test_gemm_bf16x6_general_epilogue_and_batch(chai, oracle):
    M, N, K = 200, 160, 320
    A = unif(oracle, 31, M, K)
    Bs = np.stack([unif(oracle, 32 + i, K, N) for i in range(3)])
    C0 = unif(oracle, 40, M, N)
    At, Bt = torch.from_numpy(A).cuda(), torch.from_numpy(Bs).cuda()
    C = torch.from_numpy(np.stack([C0] * 3)).cuda()
    chai.gemm(At, Bt, C=C, general=True, alpha=0.5, beta=-1.25, mode=chai.GEMM_BF16X6)
    for i in range(3):
        ref = 0.5 * (A.astype(np.float64) @ Bs[i].astype(np.float64)) - 1.25 * C0
        mag = 0.5 * (np.abs(A.astype(np.float64)) @ np.abs(Bs[i].astype(np.float64))) + 1.25 * np.abs(C0)
        assert np.all(np.abs(C[i].cpu().numpy() - ref) <= bound("bf16x6", K) * mag + U * np.abs(ref)), i
    epi = chai.epilogue(0.5, 1.0, 0.0)
    E = chai.gemm(At, Bt[0], mode=chai.GEMM_BF16X6, epi=epi).cpu().numpy()
    P = chai.gemm(At, Bt[0], mode=chai.GEMM_BF16X6).cpu().numpy()
    assert np.array_equal(E, np.maximum(P * np.float32(0.5) + np.float32(1.0), np.float32(0.0)))


@pytest.mark.parametrize("lo", [-1.0, 0.0])
@pytest.mark.parametrize("K", [256, 2048, 8192, 16384])
def test_auto_family_within_twice_fp32_error(chai, oracle, lo, K):
    """CHAI_GEMM_AUTO for f32: the family it picks keeps the error against FP64
    within twice the FP32-FMA kernel's, in max and in rms (M = N = 256, every
    element, operands uniform in [lo, 1) from Chai's stream)."""
    M = N = 256
    A = oracle.rng_fill(1230, 4, 4, 42, K, M, limits=(lo, 1.0)).reshape(M, K)
    B = oracle.rng_fill(1230, 4, 4, 42 + M * K, N, K, limits=(lo, 1.0)).reshape(K, N)
    At, Bt = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    C64 = A.astype(np.float64) @ B.astype(np.float64)
    ea = np.abs(chai.gemm(At, Bt, mode=chai.GEMM_AUTO).cpu().numpy() - C64)
    ef = np.abs(chai.gemm(At, Bt, mode=chai.GEMM_FMA).cpu().numpy() - C64)
    rms = lambda e: float(np.sqrt((e ** 2).mean()))   # noqa: E731
    print("K=%d lo=%g auto max %.3e rms %.3e fma max %.3e rms %.3e" % (K, lo, ea.max(), rms(ea), ef.max(), rms(ef)))
    assert ea.max() <= 2 * ef.max() and rms(ea) <= 2 * rms(ef)


@pytest.mark.parametrize("M,N,K", [(300, 257, 129), (1000, 64, 2048), (64, 1000, 33)])
def test_fma_tile_variants_give_the_same_bits(chai, M, N, K):
    """Every FP32-FMA tile variant forms each output as one fmaf chain over k
    in order, so all of them agree bit for bit (and the tuner compares them
    by value)."""
    g = torch.Generator(device="cpu").manual_seed(M + N + K)
    A = (torch.rand(M, K, generator=g) * 2 - 1).cuda()
    B = (torch.rand(K, N, generator=g) * 2 - 1).cuda()
    nv = len(chai.gemm_variants(chai.GEMM_FMA))
    assert nv == 4
    ref = chai.gemm(A, B, mode=chai.GEMM_FMA, variant=0)
    for v in range(1, nv):
        assert torch.equal(chai.gemm(A, B, mode=chai.GEMM_FMA, variant=v), ref), v
    err = (ref.double() - A.double() @ B.double()).abs()
    assert torch.all(err <= (K + 2) * 2.0 ** -24 * (A.double().abs() @ B.double().abs()))
