"""Benchmark: FP32 GEMM M = N = K = 8192 on one MI355X, split-BF16x3 MFMA vs
the plain FP32-FMA kernel, errors against an FP64 reference.

One "step" = one chai_hip_gemm call in split-BF16x3 mode on resident inputs
(operand split pre-pass + MFMA kernel).  A and B are uniform[-1,1] float32
from Chai's Philox4x32-10 stream (one generator, seed 42: A then B, the seed
advancing by W*H between the two calls as api/RngFun.cpp:30 does).

Prints ONE JSON line (see the fields below).  --gpus N > 1 is refused: this
build measures one GPU.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_BF16_TFLOPS = 2500.0      # MI355X dense BF16 MFMA
PEAK_FP32_TFLOPS = 157.3       # MI355X FP32 vector (incl. packed dual issue)
PEAK_HBM_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--size", type=int, default=8192)
    p.add_argument("--fp32-steps", type=int, default=3)
    p.add_argument("--cpu-size", type=int, default=1024,
                   help="edge of the interpreter-restatement sample timed on one host core")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--err-rows", type=int, default=16)
    return p.parse_args()


def cpu_baseline(n):
    """The oracle's interpreter-semantics matmul (interp/InterpMatmulBase.hpp:49-61) on one
    host core, timed on an n^3 sample of the same workload (f32 uniform[-1,1])."""
    import numpy as np
    from oracle import oracle as o
    A = o.rng_fill(1230, 4, 4, 42, n, n, limits=(-1.0, 1.0)).reshape(n, n)
    B = o.rng_fill(1230, 4, 4, 42 + n * n, n, n, limits=(-1.0, 1.0)).reshape(n, n)
    t0 = time.perf_counter()
    o.matmul_interp(A, B)
    dt = time.perf_counter() - t0
    return {"value": 2.0 * n ** 3 / dt / 1e12, "unit": "TFLOP/s", "cores": 1, "kind": "port",
            "sample": "interpreter-semantics f32 matmul %d^3 (uniform[-1,1], seed 42), 1 core, %.1f s"
                      % (n, dt)}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: --gpus %d is not supported by this build "
              "(the GEMM would shard by 2-D output tiles across GPUs; see DESIGN.md)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import chai_amd as chai

    n = args.size
    chai.init(0)
    dev = torch.device("cuda:0")
    A = torch.empty(n, n, dtype=torch.float32, device=dev)
    B = torch.empty(n, n, dtype=torch.float32, device=dev)
    chai.rng_fill(A, seed=42, limits=(-1.0, 1.0), W=n, H=n)
    chai.rng_fill(B, seed=42 + n * n, limits=(-1.0, 1.0), W=n, H=n)
    C = torch.empty(n, n, dtype=torch.float32, device=dev)
    flops = 2.0 * n * n * n

    def run(mode):
        chai.gemm(A, B, C=C, mode=mode)

    # ---- split-BF16x3: warmup, then exactly `steps` timed steps
    for _ in range(args.warmup):
        run(chai.GEMM_BF16X3)
    torch.cuda.synchronize()
    tm = chai.Timer()
    t0 = time.perf_counter()
    tm.start()
    for _ in range(args.steps):
        run(chai.GEMM_BF16X3)
    tm.stop()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_x3 = tm.ms() / args.steps
    tflops_x3 = flops / (ms_x3 * 1e-3) / 1e12
    rows = np.random.default_rng(0).choice(n, args.err_rows, replace=False)
    Cx3 = C[torch.from_numpy(rows).to(dev)].cpu().numpy().astype(np.float64)

    # ---- per-kernel durations on the launch stream (hipEvents inside the library)
    chai.profile(True)
    prep, core = [], []
    for _ in range(3):
        run(chai.GEMM_BF16X3)
        a, b = chai.gemm_last_timing()
        prep.append(a); core.append(b)
    chai.profile(False)
    core_ms = sum(core) / len(core)
    prep_ms = sum(prep) / len(prep)

    # ---- plain FP32-FMA kernel on the same inputs
    run(chai.GEMM_FMA)
    torch.cuda.synchronize()
    tm.start()
    for _ in range(args.fp32_steps):
        run(chai.GEMM_FMA)
    tm.stop()
    ms_f32 = tm.ms() / args.fp32_steps
    tflops_f32 = flops / (ms_f32 * 1e-3) / 1e12
    Cf32 = C[torch.from_numpy(rows).to(dev)].cpu().numpy().astype(np.float64)

    # ---- errors against FP64 on sampled rows (exact f32 products, float64 sums)
    Ah = A[torch.from_numpy(rows).to(dev)].cpu().numpy().astype(np.float64)
    Bh = B.cpu().numpy().astype(np.float64)
    C64 = Ah @ Bh
    e3 = np.abs(Cx3 - C64); ef = np.abs(Cf32 - C64)
    err = {"rows_sampled": int(args.err_rows),
           "bf16x3_max_abs": float(e3.max()), "bf16x3_rms": float(np.sqrt((e3 ** 2).mean())),
           "fp32_fma_max_abs": float(ef.max()), "fp32_fma_rms": float(np.sqrt((ef ** 2).mean())),
           "ratio_max": float(e3.max() / ef.max()),
           "ratio_rms": float(np.sqrt((e3 ** 2).mean()) / np.sqrt((ef ** 2).mean()))}

    mfma_flops = 3 * flops           # three bf16 products per FP32 multiply-add
    achieved = mfma_flops / (core_ms * 1e-3) / 1e12
    line = {
        "metric": "GEMM TFLOP/s (bf16x3 MFMA vs FP32) at FP32-level error; autotune cold-start s",
        "value": round(tflops_x3, 2),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_x3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (split-bf16x3 MFMA, f32 accumulate)",
        "data": "synthetic: uniform[-1,1] f32 from Chai's Philox4x32-10 stream, seed 42",
        "config": {"workload": "sgemm %dx%dx%d row-major f32, split-bf16x3 vs plain fp32-fma" % (n, n, n),
                   "M": n, "N": n, "K": n},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 1), "peak": PEAK_BF16_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_BF16_TFLOPS, 4),
                     "traffic": None,
                     "kernel": "gemm_bf16x3", "kernel_ms": round(core_ms, 4),
                     "algorithmic_flops": mfma_flops},
        "fp32_fma": {"tflops": round(tflops_f32, 2), "ms": round(ms_f32, 3),
                     "frac_of_fp32_peak": round(tflops_f32 / PEAK_FP32_TFLOPS, 4)},
        "speedup_vs_fp32": round(tflops_x3 / tflops_f32, 3),
        "split_prepass_ms": round(prep_ms, 4),
        "errors_vs_fp64": err,
        "autotune_cold_start_s": None,
        "wall_s_timed": round(wall, 3),
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.cpu_size)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
