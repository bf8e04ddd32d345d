"""GPU: the benchmark configurations at their stated sizes.

- Arrayf64 DGEMM 4096^3 on FP64 MFMA, and i32 x f32 4096^3 (ints in [-128, 127]
  from Chai's Philox words): sampled rows against the FP64 product, within the
  written bounds (2K 2^-53 sum|ab| for FP64; the split forms' bounds for the
  mixed product).
- The ragged batch: 4096 square f32 GEMMs, n log-uniform in [32, 1024] step 16
  (seed 42), epilogue max(C * 0.5 + 1, 0), submitted by 64 host threads (64
  items each) to one batcher: bit-exact against the dense path on a sample of
  instances, and within the split bound of the FP64 product on a few of them.
- The learned tile model (chai_amd/data/tile_mlp.npz, trained on autotuner
  timings of 2000 shapes) on the 500 held-out shapes of the learned-autotune
  workload (M, N, K log-uniform in [64, 16384] step 16), for each family it
  covers (BF16X3, FMA, F64 tiles and the GEMV kernels): mean throughput
  regret against the exhaustive timed search at most 5 %, cold start (one
  inference against the timed search) at least 10 times shorter.
"""
import os
import threading

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

U = 2.0 ** -24
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _uniform(chai, shape, seed, dtype=torch.float32):
    t = torch.empty(*shape, dtype=dtype, device="cuda")
    chai.rng_fill(t, seed=seed, limits=(-1.0, 1.0), W=shape[-1], H=t.numel() // shape[-1])
    return t


def _sampled_rows(A, B, C, rows):
    ridx = torch.from_numpy(rows).cuda()
    A64 = A[ridx].cpu().numpy().astype(np.float64)
    B64 = B.cpu().numpy().astype(np.float64)
    return C[ridx].cpu().numpy().astype(np.float64), A64 @ B64, np.abs(A64) @ np.abs(B64)


def test_dgemm_4096(chai):
    n = 4096
    print("dgemm 4096^3", flush=True)
    A = _uniform(chai, (n, n), 42, torch.float64)
    B = _uniform(chai, (n, n), 42 + n * n, torch.float64)
    C = chai.gemm(A, B, mode=chai.GEMM_F64)
    rows = np.random.default_rng(5).choice(n, 8, replace=False)
    got, ref, mag = _sampled_rows(A, B, C, rows)
    assert np.all(np.abs(got - ref) <= (2 * n + 2) * 2.0 ** -53 * mag)


def test_mixed_int_float_4096(chai):
    n = 4096
    print("mixed i32 x f32 4096^3", flush=True)
    W = torch.empty(n, n, dtype=torch.int32, device="cuda")
    chai.rng_fill(W, seed=42, prec=chai.U32)
    A = (W & 255) - 128                                   # lo + (w mod 256) in [-128, 127]
    B = _uniform(chai, (n, n), 42 + n * n)
    rows = np.random.default_rng(6).choice(n, 8, replace=False)
    for mode, coef in [(chai.GEMM_BF16X3, 3 * 2.0 ** -16 + 3 * n * U / (1 - 3 * n * U)),
                       (chai.GEMM_AUTO, (6 * n + 4) * U)]:
        C = chai.gemm(A, B, mode=mode, out_dtype=torch.float32)
        got, ref, mag = _sampled_rows(A, B, C, rows)
        assert np.all(np.abs(got - ref) <= coef * mag), mode


def test_ragged_4096_from_64_threads(chai, oracle):
    from chai_amd import ragged
    from chai_amd.workloads import RAGGED_EPILOGUE, ragged_sizes
    sizes = ragged_sizes()
    assert len(sizes) == 4096
    print("ragged batch: %d instances, sum 2n^3 = %.3e" % (len(sizes), float(np.sum(2.0 * sizes.astype(np.float64) ** 3))),
          flush=True)
    batch = []
    for i, k in enumerate(sizes):
        k = int(k)
        batch.append((_uniform(chai, (k, k), 42 + 2 * i), _uniform(chai, (k, k), 43 + 2 * i),
                      torch.full((k, k), float("nan"), device="cuda")))
    torch.cuda.synchronize()
    epi = chai.epilogue(*RAGGED_EPILOGUE)
    b = ragged.Batcher(len(sizes), epi=epi, variant=-1, graph=True)
    errors = []

    def worker(t):
        try:
            b.submit([ragged.item(*batch[t * 64 + j]) for j in range(64)])
        except Exception as e:   # noqa: BLE001 - reported below
            errors.append(e)

    th = [threading.Thread(target=worker, args=(t,)) for t in range(64)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    assert not errors, errors
    st = b.stats()
    assert st["batches"] == 1 and st["items"] == 4096
    b.close()
    rng = np.random.default_rng(7)
    sample = rng.choice(len(sizes), 48, replace=False)
    for i in sample:
        A, B, C = batch[i]
        # every tile variant gives the same C: the dense path's variant 2
        assert torch.equal(C, chai.gemm(A, B, mode=chai.GEMM_BF16X3, variant=2, epi=epi)), i
    for i in sample[:4]:
        A, B, C = batch[i]
        k = A.shape[0]
        A64, B64 = A.cpu().numpy().astype(np.float64), B.cpu().numpy().astype(np.float64)
        P64, mag = A64 @ B64, np.abs(A64) @ np.abs(B64)
        pre = P64 * 0.5 + 1.0
        # the epilogue adds at most two f32 roundings on top of the product's bound
        tol = 0.5 * (3 * 2.0 ** -16 + 3 * k * U / (1 - 3 * k * U)) * mag + 4 * U * np.abs(pre) + 1e-30
        assert np.all(np.abs(C.cpu().numpy() - np.maximum(pre, 0.0)) <= tol), i


def test_learned_tuner_on_held_out_config_shapes(chai):
    from chai_amd import tuner
    model_file = os.path.join(ROOT, "chai_amd", "data", "tile_mlp.npz")
    m = tuner.load_model(model_file)
    ws = tuner.Workspace(hi=16384)
    test = tuner.random_shapes(500, seed=2025)
    for f in tuner.FAMILIES:
        name = tuner.FAMILY_NAMES[f]
        tuner.predict(m, test[0], f)             # one-time module setup
        times, cold_ex, choice, cold_m = [], [], [], []
        for i, s in enumerate(test):
            ms, secs = tuner.time_variants(ws, s, trials=2, family=f)
            times.append(ms)
            cold_ex.append(secs)
            v, _, secs_m = tuner.predict(m, s, f)
            choice.append(v)
            cold_m.append(secs_m)
            if i % 100 == 0:
                print("%s: held-out shape %d/500" % (name, i), flush=True)
        times = np.array(times)
        assert np.all(times > 0)
        reg = tuner.regret(choice, times)
        speed = np.mean(cold_ex) / np.mean(cold_m)
        print("%s: mean regret %.4f, median %.4f, p90 %.4f; cold start %.2e s vs %.2e s (%.0fx)"
              % (name, reg.mean(), np.median(reg), np.quantile(reg, 0.9), np.mean(cold_m), np.mean(cold_ex), speed),
              flush=True)
        assert reg.mean() <= 0.05, name
        assert speed >= 10.0, name
