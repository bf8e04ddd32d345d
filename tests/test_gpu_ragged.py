"""GPU parity of ragged batches (chai_hip_ragged_*, chai_hip_batcher_*).

The persistent kernel runs the same split-BF16x3 tile arithmetic as the dense
path, so a ragged batch must reproduce per-instance chai_hip_gemm results
BIT-EXACTLY (same k order, same MFMAs), with the fused epilogue equal to the
f32 ops C*scale, +shift, max(.., lo) applied one by one.  Against the FP64
oracle the BF16X3 bound applies: |C - C64| <= (3 * 2^-16 + gamma_3K) sum|a b|.
"""
import threading

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

U = 2.0 ** -24
EPI = (0.5, 1.0, 0.0)     # max(C * 0.5f + 1, 0)


def unif(oracle, seed, rows, cols):
    return oracle.rng_fill(1230, 4, 0, seed, cols, rows, 1, limits=(-1.0, 1.0)).reshape(rows, cols)


def make_batch(oracle, shapes, seed0=100):
    out = []
    for i, (M, N, K) in enumerate(shapes):
        A = torch.from_numpy(unif(oracle, seed0 + 2 * i, M, K)).cuda()
        B = torch.from_numpy(unif(oracle, seed0 + 2 * i + 1, K, N)).cuda()
        C = torch.full((M, N), float("nan"), device="cuda")
        out.append((A, B, C))
    return out


def epi_ref(C):
    s, h, lo = EPI
    t = C * torch.tensor(s, dtype=torch.float32, device=C.device)   # f32 multiply
    t = t + torch.tensor(h, dtype=torch.float32, device=C.device)   # f32 add
    return torch.clamp_min(t, lo)


SHAPES = [(1, 1, 1), (17, 33, 65), (300, 129, 1000), (64, 64, 64), (1024, 96, 48), (33, 512, 7),
          (256, 256, 256), (129, 257, 513), (16, 1024, 16), (500, 500, 500)]


@pytest.mark.parametrize("variant", [0, 2, 5, 7, 8, 9])
@pytest.mark.parametrize("with_epi", [False, True])
def test_ragged_matches_dense_bit_exact(chai, oracle, variant, with_epi):
    """Variants 0..7 run the dense path's panel tiles after a split pre-pass;
    8 and 9 read the f32 operands directly and split inside the tile.  Every
    tile shape accumulates each output in the same k order with the same
    MFMAs, so all of them equal the dense GEMM bit for bit."""
    from chai_amd import ragged
    print("case ragged variant", variant, "epi", with_epi, flush=True)
    batch = make_batch(oracle, SHAPES)
    epi = chai.epilogue(*EPI) if with_epi else None
    plan = ragged.Plan([ragged.item(A, B, C) for A, B, C in batch], epi=epi, variant=variant)
    info = plan.info()
    assert info["variant"] == variant and info["tiles"] > 0
    assert (info["split_blocks"] == 0) == (variant >= 8)
    plan.run()
    torch.cuda.synchronize()
    dv = variant if variant < 8 else 2
    for (A, B, C), (M, N, K) in zip(batch, SHAPES):
        D = chai.gemm(A, B, mode=chai.GEMM_BF16X3, variant=dv)
        if with_epi:
            assert torch.equal(C, epi_ref(D)), (M, N, K)
            E = chai.gemm(A, B, mode=chai.GEMM_BF16X3, variant=dv, epi=epi)
            assert torch.equal(C, E), (M, N, K)
        else:
            assert torch.equal(C, D), (M, N, K)
    plan.release()


@pytest.mark.parametrize("variant", [8, 9])
def test_ragged_f32_fed_strided_and_graph(chai, oracle, variant):
    """The f32-fed tiles on sub-matrix views (unaligned rows, odd strides) and
    through graph replays, against the dense GEMM of contiguous copies."""
    from chai_amd import ragged
    big = torch.from_numpy(unif(oracle, 77, 700, 900)).cuda()
    views = [(big[3:300, 5:262], big[1:258, 400:517]),      # odd offsets: scalar loads
             (big[100:164, 0:64], big[200:264, 512:704]),   # aligned rows
             (big[10:11, 7:8], big[20:21, 30:31])]          # 1 x 1 x 1
    items, outs = [], []
    for A, B in views:
        C = torch.full((A.shape[0], B.shape[1]), float("nan"), device="cuda")
        outs.append(C)
        items.append(ragged.item(A, B, C))
    plan = ragged.Plan(items, variant=variant)
    for graph in (False, True, True):
        for C in outs:
            C.fill_(float("nan"))
        torch.cuda.synchronize()
        plan.run(graph=graph)
        torch.cuda.synchronize()
        for (A, B), C in zip(views, outs):
            assert torch.equal(C, chai.gemm(A.contiguous(), B.contiguous(), mode=chai.GEMM_BF16X3)), (A.shape, graph)
    plan.release()


def test_ragged_vs_fp64_oracle(chai, oracle):
    from chai_amd import ragged
    rng = np.random.default_rng(42)
    shapes = [(int(n), int(n), int(n)) for n in
              (np.exp(rng.uniform(np.log(32), np.log(1024), 40)) // 16 * 16).astype(int)]
    batch = make_batch(oracle, shapes, seed0=7000)
    plan = ragged.Plan([ragged.item(A, B, C) for A, B, C in batch])
    plan.run()
    torch.cuda.synchronize()
    worst = 0.0
    for (A, B, C), (M, N, K) in zip(batch, SHAPES):
        A64 = A.cpu().double().numpy(); B64 = B.cpu().double().numpy()
        C64 = A64 @ B64
        absprod = np.abs(A64) @ np.abs(B64)
        g = K * U / (1 - K * U)
        err = np.abs(C.cpu().double().numpy() - C64)
        g3 = 3 * K * U / (1 - 3 * K * U)
        assert np.all(err <= (3 * 2.0 ** -16 + g3) * absprod), (M, N, K)
        worst = max(worst, float((err / absprod).max()))
    print("max |C - C64| / sum|ab| =", worst)


def test_ragged_graph_replays(chai, oracle):
    """Captured hipGraph: replays with eager work in between must all equal the eager run."""
    from chai_amd import ragged
    batch = make_batch(oracle, SHAPES[:6], seed0=300)
    epi = chai.epilogue(*EPI)
    plan = ragged.Plan([ragged.item(A, B, C) for A, B, C in batch], epi=epi)
    plan.run()
    torch.cuda.synchronize()
    ref = [C.clone() for _, _, C in batch]
    for rep in range(4):
        for _, _, C in batch:
            C.fill_(float(rep) - 7.0)                # eager work between replays
        torch.cuda.synchronize()
        plan.run(graph=True)
        torch.cuda.synchronize()
        for (_, _, C), R in zip(batch, ref):
            assert torch.equal(C, R), rep
    # new inputs flow through the same graph
    A0, B0, C0 = batch[0]
    A0.mul_(2.0)
    plan.run(graph=True)
    torch.cuda.synchronize()
    assert torch.equal(C0, epi_ref(chai.gemm(A0, B0, mode=chai.GEMM_BF16X3)))


def test_ragged_edge_cases(chai, oracle):
    from chai_amd import ragged
    # empty batch and empty outputs are no-ops
    p = ragged.Plan([])
    p.run()
    p.run(graph=True)
    assert p.info()["tiles"] == 0
    A = torch.ones(0, 5, device="cuda"); B = torch.ones(5, 3, device="cuda"); C = torch.ones(0, 3, device="cuda")
    p = ragged.Plan([ragged.item(A, B, C)])
    p.run()
    torch.cuda.synchronize()
    # strided (sub-matrix) operands
    big = torch.from_numpy(unif(oracle, 5, 200, 300)).cuda()
    A = big[10:110, 20:70]; B = big[:50, 100:180]; C = torch.empty(100, 200, device="cuda")[:, 50:130]
    p = ragged.Plan([ragged.item(A, B, C)])
    p.run()
    torch.cuda.synchronize()
    assert torch.equal(C, chai.gemm(A.contiguous(), B.contiguous(), mode=chai.GEMM_BF16X3))
    # bad descriptors are rejected
    it = ragged.item(A, B, C)
    it.lda = 10
    with pytest.raises(chai.ChaiError):
        ragged.Plan([it])
    with pytest.raises(chai.ChaiError):
        ragged.Plan([ragged.item(A, B, C)], variant=99)


def test_ragged_f32_fed_row_stride_limit(chai, oracle):
    """The f32-fed tiles address rows through 31-bit buffer offsets: a row
    stride of 2^23 floats is refused for them, accepted by the panel tiles."""
    from chai_amd import ragged
    big = torch.zeros(2 * (1 << 23) + 64, device="cuda")
    A = big.as_strided((2, 16), (1 << 23, 1))
    B = torch.ones(16, 16, device="cuda")
    C = torch.empty(2, 16, device="cuda")
    for v in (8, 9):
        with pytest.raises(chai.ChaiError):
            ragged.Plan([ragged.item(A, B, C)], variant=v)
    p = ragged.Plan([ragged.item(A, B, C)], variant=2)
    p.run()
    torch.cuda.synchronize()
    assert torch.equal(C, torch.zeros(2, 16, device="cuda"))
    p.release()


def test_ragged_tune_picks_a_variant(chai, oracle):
    from chai_amd import ragged
    batch = make_batch(oracle, SHAPES, seed0=900)
    items = [ragged.item(A, B, C) for A, B, C in batch]
    v, ms = ragged.tune(items, trials=1)
    assert ragged.num_variants() == 10
    assert 0 <= v < ragged.num_variants() and ms > 0
    p = ragged.Plan(items)                     # variant -1 -> the tuned one
    assert p.info()["variant"] == v


def test_batcher_many_threads(chai, oracle):
    """16 host threads x 8 items each gathered into one batch of 128; a second
    round with the same items replays the graph."""
    from chai_amd import ragged
    rng = np.random.default_rng(3)
    nthreads, per = 16, 8
    shapes = [(int(n), int(n), int(n)) for n in
              (np.exp(rng.uniform(np.log(32), np.log(384), nthreads * per)) // 16 * 16).astype(int)]
    batch = make_batch(oracle, shapes, seed0=2000)
    epi = chai.epilogue(*EPI)
    b = ragged.Batcher(nthreads * per, epi=epi, variant=2, graph=True)
    errors = []

    def worker(t):
        try:
            b.submit([ragged.item(*batch[t * per + j]) for j in range(per)])
        except Exception as e:   # noqa: BLE001 - reported below
            errors.append(e)

    for rnd in range(2):
        for _, _, C in batch:
            C.fill_(float("nan"))
        torch.cuda.synchronize()
        th = [threading.Thread(target=worker, args=(t,)) for t in range(nthreads)]
        for x in th:
            x.start()
        for x in th:
            x.join()
        assert not errors, errors
        for (A, B, C) in batch:
            assert torch.equal(C, chai.gemm(A, B, mode=chai.GEMM_BF16X3, variant=2, epi=epi))
    st = b.stats()
    assert st["batches"] == 2 and st["items"] == 2 * nthreads * per
    assert st["plans"] == 1 and st["replays"] == 1
    b.close()


def test_batcher_timeout_runs_partial_batch(chai, oracle):
    from chai_amd import ragged
    batch = make_batch(oracle, [(64, 64, 64), (100, 50, 20)], seed0=4000)
    b = ragged.Batcher(1000, variant=2, graph=False, timeout_ms=20)
    b.submit([ragged.item(*x) for x in batch])       # never fills: runs after the timeout
    for A, B, C in batch:
        assert torch.equal(C, chai.gemm(A, B, mode=chai.GEMM_BF16X3, variant=2))
    assert b.stats()["batches"] == 1
    b.close()


def test_batcher_waits_for_the_producer_stream(chai, oracle):
    """Submitters write their operands on their own streams and submit at once,
    without a host synchronize: the batch must run after that work.  The copy
    into A queues behind a spin kernel, so a batch that did not wait would
    read the NaN that A holds before the copy."""
    from chai_amd import ragged
    nthreads = 4
    batch = make_batch(oracle, [(96, 80, 64), (128, 128, 128), (64, 200, 48), (256, 64, 96)], seed0=5000)
    src = [A.clone() for A, _, _ in batch]
    for A, _, C in batch:
        A.fill_(float("nan"))
        C.fill_(float("nan"))
    torch.cuda.synchronize()
    b = ragged.Batcher(nthreads, variant=2, graph=False)
    errors = []

    def worker(t):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                torch.cuda._sleep(20_000_000)          # tens of ms of device time before the copy
                batch[t][0].copy_(src[t])
            b.submit([ragged.item(*batch[t])], stream=s)
        except Exception as e:   # noqa: BLE001 - reported below
            errors.append(e)

    th = [threading.Thread(target=worker, args=(t,)) for t in range(nthreads)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    assert not errors, errors
    torch.cuda.synchronize()
    for (A, B, C), s in zip(batch, src):
        assert torch.equal(A, s)
        assert torch.equal(C, chai.gemm(A, B, mode=chai.GEMM_BF16X3, variant=2))
    b.close()
