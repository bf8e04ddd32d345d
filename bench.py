"""Benchmarks of the MI355X back end; each run prints ONE JSON line.

Default workload (what a plain `python bench.py` measures): FP32 GEMM
M = N = K = 8192 on one MI355X, split-BF16x3 MFMA vs the plain FP32-FMA kernel,
errors against an FP64 reference; then, in the same run and line ("legs"), the
other configurations each with a roofline and a CPU baseline (--no-legs skips
them).  One "step" = one chai_hip_gemm call in
split-BF16x3 mode on resident inputs (operand split pre-pass + MFMA kernel).
A and B are uniform[-1,1] float32 from Chai's Philox4x32-10 stream (one
generator, seed 42: A then B, the seed advancing by W*H between the two calls
as api/RngFun.cpp:30 does).

Other workloads (--workload): ragged (4096-instance ragged GEMM batch + fused
epilogue), autotune (learned tile model: regret and cold start), mc (fused
Monte Carlo trace, Philox normals, GEMV 16384^2), dgemm (FP64 MFMA 4096^3),
mixed (i32 x f32 4096^3), gather (gather2_floor stencil 8192^2).  --profile-run is the short program profiled with
rocprofv3; --pmc-summary turns two counter runs into the traffic summary.

--gpus N > 1 is refused: this build measures one GPU.
"""
import argparse
import csv
import glob
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_BF16_TFLOPS = 2500.0      # MI355X dense BF16 MFMA
PEAK_FP32_TFLOPS = 157.3       # MI355X FP32 vector
PEAK_FP64_TFLOPS = 78.6        # MI355X FP64 MFMA
PEAK_HBM_GBS = 8000.0
METRIC = "GEMM TFLOP/s (bf16x3 MFMA vs FP32) at FP32-level error; autotune cold-start s"
PMC_SUMMARY = os.path.join(ROOT, "profiles", "gemm_bf16x3_8192_pmc.txt")
PMC_SUMMARY_GATHER = os.path.join(ROOT, "profiles", "gather_stencil_8192_pmc.txt")
PMC_SUMMARY_STREAM = os.path.join(ROOT, "profiles", "stream_trace_2e28_pmc.txt")
MODEL_FILE = os.path.join(ROOT, "chai_amd", "data", "tile_mlp.npz")


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--workload", default="sgemm",
                   choices=["sgemm", "ragged", "autotune", "mc", "dgemm", "mixed", "gather", "errors", "stream"])
    p.add_argument("--size", type=int, default=8192)
    p.add_argument("--fp32-steps", type=int, default=3)
    p.add_argument("--cpu-size", type=int, default=1024,
                   help="edge of the interpreter-restatement sample timed on one host core")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--cpu-budget", type=float, default=25.0,
                   help="seconds of host work for a sampled CPU baseline")
    p.add_argument("--no-legs", action="store_true",
                   help="default workload: only the headline GEMM, not the other configurations' legs")
    p.add_argument("--err-rows", type=int, default=16)
    p.add_argument("--train-shapes", type=int, default=2000)
    p.add_argument("--test-shapes", type=int, default=500)
    p.add_argument("--save-model", default=None, help="autotune: write the trained model here")
    p.add_argument("--profile-run", action="store_true",
                   help="short program for rocprofv3: sgemm (3 GEMMs plus the calibration kernels) or, "
                        "with --workload ragged, the ragged batch from two operand buffers")
    p.add_argument("--ragged-variant", type=int, default=2, help="tile variant of the ragged profile run")
    p.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                   help="summarise two rocprofv3 --pmc runs of --profile-run (no GPU needed)")
    p.add_argument("--pmc-table", nargs="+", metavar="DIR",
                   help="per-kernel counter values of rocprofv3 --pmc runs (no GPU needed)")
    p.add_argument("--kernel-stats", nargs=3, metavar=("DIR", "TITLE", "COMMAND"),
                   help="summarise a rocprofv3 --kernel-trace --stats run (no GPU needed)")
    return p.parse_args()


def emit(line):
    print(json.dumps(line), flush=True)


def base_line(args, value, unit, ms, dtype, config, higher=True):
    return {"metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": higher, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype,
            "data": "synthetic: uniform[-1,1] from Chai's Philox4x32-10 stream", "config": config}


def timed(fn, steps, warmup):
    """hipEvent time (ms) per call of fn() on torch's current stream"""
    import torch
    import chai_amd as chai
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    tm = chai.Timer()
    tm.start()
    for _ in range(steps):
        fn()
    tm.stop()
    torch.cuda.synchronize()
    return tm.ms() / steps


def uniform(shape, seed, dtype=None):
    import torch
    import chai_amd as chai
    t = torch.empty(*shape, dtype=dtype or torch.float32, device="cuda")
    chai.rng_fill(t, seed=seed, limits=(-1.0, 1.0), W=shape[-1], H=t.numel() // shape[-1])
    return t


# ------------------------------------------------------------------ sgemm
def cpu_baseline_gemm(n):
    """The oracle's interpreter-semantics matmul (interp/InterpMatmulBase.hpp:49-61) on one
    host core, timed on an n^3 sample of the same workload (f32 uniform[-1,1])."""
    from oracle import oracle as o
    A = o.rng_fill(1230, 4, 4, 42, n, n, limits=(-1.0, 1.0)).reshape(n, n)
    B = o.rng_fill(1230, 4, 4, 42 + n * n, n, n, limits=(-1.0, 1.0)).reshape(n, n)
    t0 = time.perf_counter()
    o.matmul_interp(A, B)
    dt = time.perf_counter() - t0
    out = {"value": 2.0 * n ** 3 / dt / 1e12, "unit": "TFLOP/s", "cores": 1, "kind": "port",
           "sample": "the oracle's restatement of the interpreter: f32 matmul %d^3 (uniform[-1,1], seed 42), "
                     "1 core, %.1f s" % (n, dt)}
    out.update(host_info())
    return out


def host_info():
    """nproc and CPU model of the machine the CPU baseline ran on"""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for l in f:
                if l.startswith("model name"):
                    model = l.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def cpu_reference_gemm(n):
    """The reference's own interpreter (compiled from its sources into
    oracle/_ref by oracle/frontend.mk and tests/frontend/Makefile) timed on one
    host core: C = matmul(A, B) on n x n f32 through chai.h, from the
    statements to the readout that evaluates them.  None when not built."""
    import subprocess
    exe = os.path.join(ROOT, "oracle", "_ref", "fe_bin", "fe_cases")
    if not os.path.exists(exe):
        return None
    env = dict(os.environ, OCL_ICD_VENDORS="/nonexistent")
    env.pop("CHAI_MI355X", None)
    try:
        r = subprocess.run([exe, "mm_time", str(n)], env=env, capture_output=True, text=True, timeout=300)
        m = re.search(r"seconds ([0-9.eE+-]+) flops ([0-9.eE+-]+)", r.stdout)
    except (OSError, subprocess.SubprocessError):
        return None
    if r.returncode != 0 or not m:
        return None
    dt, fl = float(m.group(1)), float(m.group(2))
    out = {"value": fl / dt / 1e12, "unit": "TFLOP/s", "cores": 1, "kind": "reference",
           "sample": "the reference's interpreter (built from its sources), f32 matmul %d^3 through chai.h, "
                     "1 core, %.1f s" % (n, dt)}
    out.update(host_info())
    return out


def read_traffic(path=PMC_SUMMARY):
    """HBM bytes per launch of the measured kernel from a committed counter summary."""
    try:
        with open(path) as f:
            m = re.search(r"traffic_bytes_per_launch\s*=\s*([0-9.eE+]+)", f.read())
        return float(m.group(1)) if m else None
    except OSError:
        return None


def cold_start(A, B, C):
    """Wall seconds from a new shape to a chosen tile variant: exhaustive timed
    search vs one inference of the shipped tile model (None without a model)."""
    import numpy as np
    import torch
    import chai_amd as chai
    from chai_amd import tuner
    n = A.shape[0]
    shape = (n, n, n, 0, 0, chai.F32)
    out = {}
    if os.path.exists(MODEL_FILE):
        m = tuner.load_model(MODEL_FILE)
        tuner.predict(m, shape)        # first call pays one-time setup (module load)
        v, _, secs = tuner.predict(m, shape)
        out.update({"model_s": round(secs, 6), "model_variant": v})
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bv, _ = chai.gemm_tune(A, B, C, trials=3)
    out.update({"exhaustive_s": round(time.perf_counter() - t0, 4), "exhaustive_variant": bv})
    if "model_s" in out:
        out["speedup"] = round(out["exhaustive_s"] / out["model_s"], 1)
    return out


def profile_run_ragged(args):
    """Short program for rocprofv3 counter runs of the ragged batch: the 4096
    instances as views of two large operand buffers (two fills instead of
    8192), one plan, three runs."""
    import torch
    import chai_amd as chai
    from chai_amd import ragged
    from chai_amd.workloads import RAGGED_EPILOGUE, ragged_sizes
    chai.init(0)
    sizes = [int(k) for k in ragged_sizes()]
    tot = sum(k * k for k in sizes)
    bufA = torch.empty(tot, device="cuda"); bufB = torch.empty(tot, device="cuda")
    bufC = torch.empty(tot, device="cuda")
    chai.rng_fill(bufA, seed=42, limits=(-1.0, 1.0))
    chai.rng_fill(bufB, seed=43, limits=(-1.0, 1.0))
    items, off = [], 0
    for k in sizes:
        v = lambda b: b[off:off + k * k].view(k, k)   # noqa: E731
        items.append(ragged.item(v(bufA), v(bufB), v(bufC)))
        off += k * k
    plan = ragged.Plan(items, epi=chai.epilogue(*RAGGED_EPILOGUE), variant=args.ragged_variant)
    for _ in range(3):
        plan.run()
    torch.cuda.synchronize()
    print("ragged profile run done", plan.info(), flush=True)


def profile_run(args):
    """Short program for rocprofv3: calibration kernels of known traffic, then
    the tile search and three bf16x3 GEMMs (8192^3) with the variant it picks,
    then the same for bf16x6."""
    import torch
    import chai_amd as chai
    chai.init(0)
    n = args.size
    big = torch.empty(1 << 28, dtype=torch.float32, device="cuda")     # 1 GiB
    chai.rng_fill(big, seed=1)            # writes exactly 2^30 bytes
    chai.reduce(big, "sum")               # reads exactly 2^30 bytes
    torch.empty_like(big).copy_(big)      # plain device copy: reads and writes 2^30 bytes
    A, B = uniform((n, n), 42), uniform((n, n), 42 + n * n)
    C = torch.empty(n, n, device="cuda")
    chai.gemm_tune(A, B, C, trials=1)     # the variant the benchmark runs
    for _ in range(3):
        chai.gemm(A, B, C=C, mode=chai.GEMM_BF16X3)
    chai.gemm_tune(A, B, C, mode=chai.GEMM_BF16X6, trials=1)
    for _ in range(3):
        chai.gemm(A, B, C=C, mode=chai.GEMM_BF16X6)
    torch.cuda.synchronize()
    print("profile run done", flush=True)


def stream_trace(n):
    """C = max(A * 0.5f + B, 0) over n f32: (trace, A, B, C)"""
    from chai_amd import F32
    from chai_amd import trace as tr
    t = tr.Trace()
    A, B = t.input(F32, n), t.input(F32, n)
    C = t.output(F32, n)
    t.assign(C, tr.maxe(A * tr.f32(0.5) + B, tr.f32(0.0)))
    return t, A, B, C


def profile_run_gather(args):
    """Short program for rocprofv3: the calibration kernels, then three JIT
    launches of the 8192^2 gather stencil (or, for --workload stream, of the
    2^28-element elementwise trace)."""
    import torch
    import chai_amd as chai
    from chai_amd.workloads import filter_trace
    chai.init(0)
    big = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    chai.rng_fill(big, seed=1)
    chai.reduce(big, "sum")
    torch.empty_like(big).copy_(big)      # plain device copy: reads and writes 2^30 bytes
    del big
    if args.workload == "stream":
        n = 1 << 28
        t, A, B, C = stream_trace(n)
        bufs = {A: uniform((n,), 42), B: uniform((n,), 43), C: torch.empty(n, device="cuda")}
    else:
        N = 8192
        t, A, B = filter_trace(N)
        bufs = {A: uniform((N, N), 42).view(-1), B: torch.empty(N * N, device="cuda")}
    prog = t.compile()
    for _ in range(3):
        prog.run(bufs)
    torch.cuda.synchronize()
    print("profile run done", flush=True)


def _counter_rows(d, counter):
    """{kernel name: [value per dispatch]} from a rocprofv3 output directory
    (SQLite database, or CSV when written with --output-format csv)."""
    rows = {}
    for fn in glob.glob(os.path.join(d, "**", "*.db"), recursive=True):
        import sqlite3
        con = sqlite3.connect(fn)
        for k, v in con.execute("select kernel_name, value from counters_collection where counter_name = ?",
                                (counter,)):
            rows.setdefault(k, []).append(float(v))
        con.close()
    for fn in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(fn) as f:
            for r in csv.DictReader(f):
                if r.get("Counter_Name") == counter:
                    rows.setdefault(r.get("Kernel_Name", "?"), []).append(float(r["Counter_Value"]))
    return rows


def pmc_summary(fetch_dir, write_dir, workload="sgemm"):
    """Per-kernel FETCH_SIZE / WRITE_SIZE (KiB per dispatch) from two rocprofv3 runs of
    --profile-run, each scaled by what the same counter reads for a plain copy
    kernel of known size (the 2^30-byte device copy: the copy dispatch with
    the largest count); the readings of rng_fill (writes 2^30 B) and of the
    reduction's partial kernel (reads 2^30 B) are listed beside it."""
    fe, wr = _counter_rows(fetch_dir, "FETCH_SIZE"), _counter_rows(write_dir, "WRITE_SIZE")
    f_copy = max([v for k, r in fe.items() if "copy" in k.lower() for v in r] or [0.0]) or None
    w_copy = max([v for k, r in wr.items() if "copy" in k.lower() for v in r] or [0.0]) or None

    def pick(rows, pat):
        # the matching kernel with the most dispatches (the tile search runs
        # every variant; the GEMMs after it repeat the chosen one)
        best = None
        for k, v in rows.items():
            if re.search(pat, k) and (best is None or len(v) > len(rows[best])):
                best = k
        return (best, sum(rows[best]) / len(rows[best])) if best else (None, None)
    _, f_cal = pick(fe, r"reduce_partial")
    _, w_cal = pick(wr, r"rng_fill_32<float, 1230, 4, 0, 0>")   # the unit fill of 2^30 B (not the [-1, 1) operand fills)
    trace_wl = workload in ("gather", "stream")
    pat = r"chai_trace_kernel" if trace_wl else r"gemm_x3_kernel"
    kf, f_gemm = pick(fe, pat)
    kw, w_gemm = pick(wr, pat)
    fetch_scale = (2 ** 30) / (f_copy * 1024) if f_copy else None
    write_scale = (2 ** 30) / (w_copy * 1024) if w_copy else None
    wl = " --workload " + workload if trace_wl else ""
    title = {"gather": "gather2_floor 5-point stencil 8192^2 f32 (JIT kernel)",
             "stream": "fused elementwise trace C = max(A*0.5f + B, 0), 2^28 f32 (JIT kernel)"}
    lines = ["# %s: HBM traffic per launch from rocprofv3 counters" % title.get(workload, "bf16x3 GEMM 8192^3"),
             "# commands (from the repository root):",
             "#   rocprofv3 --pmc FETCH_SIZE -d <fetch_dir> -o run -- python3 bench.py --profile-run" + wl,
             "#   rocprofv3 --pmc WRITE_SIZE -d <write_dir> -o run -- python3 bench.py --profile-run" + wl,
             "#   python3 bench.py --pmc-summary <fetch_dir> <write_dir>" + wl,
             "calibration_fetch_KiB(device copy, 2^30 B read) = %s" % f_copy,
             "calibration_write_KiB(device copy, 2^30 B written) = %s" % w_copy,
             "reading_fetch_KiB(reduce_partial, 2^30 B read) = %s" % f_cal,
             "reading_write_KiB(rng_fill_32, 2^30 B written) = %s" % w_cal,
             "fetch_scale = %s" % fetch_scale, "write_scale = %s" % write_scale,
             "kernel = %s" % (kf or kw),
             "fetch_KiB_per_dispatch = %s" % f_gemm, "write_KiB_per_dispatch = %s" % w_gemm]
    if f_gemm and w_gemm and fetch_scale and write_scale:
        t = f_gemm * 1024 * fetch_scale + w_gemm * 1024 * write_scale
        n = 8192
        alg = {"gather": 8 * n * n, "stream": 12 * 2 ** 28}.get(workload, 2 * 2 * 2 * n * n + 4 * n * n)
        lines.append("traffic_bytes_per_launch = %.0f" % t)
        lines.append("algorithmic_bytes_per_launch = %d  # %s" % (alg, {
            "gather": "A read once, B written once (f32)",
            "stream": "A and B read once, C written once (f32)"}.get(
                workload, "bf16 hi+lo panels of A and B^T read once, C written")))
        lines.append("traffic_over_algorithmic = %.2f" % (t / alg))
    lines.append("# per kernel, mean per dispatch (KiB, unscaled):")
    for k in sorted(set(fe) | set(wr)):
        lines.append("  %-70.70s fetch %-12s write %s" % (
            k, "%.0f" % (sum(fe[k]) / len(fe[k])) if k in fe else "-",
            "%.0f" % (sum(wr[k]) / len(wr[k])) if k in wr else "-"))
    print("\n".join(lines))


def kernel_stats(d, title, command):
    """Per-kernel call counts and durations from a rocprofv3 --kernel-trace --stats run."""
    import sqlite3
    lines = ["# %s" % title, "# command (from the repository root):", "#   %s" % command,
             "# kernel | calls | total ms | average us | percent"]
    for fn in glob.glob(os.path.join(d, "**", "*.db"), recursive=True):
        con = sqlite3.connect(fn)
        for name, calls, total, avg, pct in con.execute(
                "select name, total_calls, total_duration, average, percentage from top_kernels"):
            lines.append("%-100.100s | %5d | %10.3f | %10.3f | %5.2f" % (name, calls, total / 1e3, avg, pct))
        con.close()
    for fn in glob.glob(os.path.join(d, "**", "*kernel_stats.csv"), recursive=True):
        with open(fn) as f:
            for r in csv.DictReader(f):
                lines.append("%-100.100s | %5s | %10.3f | %10.3f | %5s" % (
                    r["Name"], r["Calls"], float(r["TotalDurationNs"]) / 1e6, float(r["AverageNs"]) / 1e3,
                    r.get("Percentage", "")))
    # one kernel name can run several shapes (the 8192^3 and 4096^3 GEMMs share
    # a tile): durations per launch grid, the median beside the mean
    shapes = {}
    for fn in glob.glob(os.path.join(d, "**", "*.db"), recursive=True):
        con = sqlite3.connect(fn)
        for name, dur, gx, gy, gz in con.execute("select name, duration, grid_x, grid_y, grid_z from kernels"):
            shapes.setdefault((name, (gx, gy, gz)), []).append(dur / 1e3)
        con.close()
    if shapes:
        lines.append("# per launch grid (work-items x, y, z), the 30 longest by total time: "
                     "kernel | grid | calls | mean us | median us | min us")
        top = sorted(shapes.items(), key=lambda kv: -sum(kv[1]))[:30]
        for (name, g), v in top:
            v = sorted(v)
            lines.append("%-80.80s | %s | %5d | %10.3f | %10.3f | %10.3f" % (
                name, "x".join(str(x) for x in g), len(v), sum(v) / len(v), v[len(v) // 2], v[0]))
    print("\n".join(lines))


def pmc_table(dirs):
    """Mean per-dispatch value of every counter collected in rocprofv3 --pmc runs
    (one counter per run), per kernel."""
    import sqlite3
    vals = {}
    for d in dirs:
        for fn in glob.glob(os.path.join(d, "**", "*.db"), recursive=True):
            con = sqlite3.connect(fn)
            for k, c, v in con.execute("select kernel_name, counter_name, avg(value) from counters_collection "
                                       "group by kernel_name, counter_name"):
                vals.setdefault(k, {})[c] = v
            con.close()
    lines = []
    for k in sorted(vals):
        lines.append(k[:110])
        for c in sorted(vals[k]):
            lines.append("    %-32s %.6g" % (c, vals[k][c]))
    print("\n".join(lines))


def run_sgemm(args):
    import numpy as np
    import torch
    import chai_amd as chai
    n = args.size
    chai.init(0)
    dev = torch.device("cuda:0")
    A, B = uniform((n, n), 42), uniform((n, n), 42 + n * n)
    C = torch.empty(n, n, dtype=torch.float32, device=dev)
    flops = 2.0 * n * n * n

    def run(mode):
        chai.gemm(A, B, C=C, mode=mode)

    # autotune first (as the reference does on a new exogenous key): the model's
    # cold-start prediction vs the exhaustive timed search, whose choice the
    # timed GEMMs below then use
    cold = cold_start(A, B, C)

    # split-BF16x3: warmup, then exactly `steps` timed steps (hipEvents on the launch stream);
    # the library also records events around each step's pre-pass and core kernel
    # (no host waits between steps), read after the timed region
    t0 = time.perf_counter()
    chai.profile(True)
    ms_x3 = timed(lambda: run(chai.GEMM_BF16X3), args.steps, args.warmup)
    prep_ms, core_ms = chai.gemm_profile_mean(min(args.steps, 64))
    chai.profile(False)
    wall = time.perf_counter() - t0
    tflops_x3 = flops / (ms_x3 * 1e-3) / 1e12
    rows = np.random.default_rng(0).choice(n, args.err_rows, replace=False)
    ridx = torch.from_numpy(rows).to(dev)
    Cx3 = C[ridx].cpu().numpy().astype(np.float64)

    # split-BF16x6 (three pieces, six products): FP32-FMA-level error at twice the MFMA work;
    # its own tile search first, then timed like BF16x3 (core kernel from the library's events)
    v6, _ = chai.gemm_tune(A, B, C, mode=chai.GEMM_BF16X6, trials=2)
    chai.profile(True)
    ms_x6 = timed(lambda: run(chai.GEMM_BF16X6), args.fp32_steps, 1)
    prep6_ms, core6_ms = chai.gemm_profile_mean(min(args.fp32_steps, 64))
    chai.profile(False)
    tflops_x6 = flops / (ms_x6 * 1e-3) / 1e12
    Cx6 = C[ridx].cpu().numpy().astype(np.float64)

    ms_f32 = timed(lambda: run(chai.GEMM_FMA), args.fp32_steps, 1)
    tflops_f32 = flops / (ms_f32 * 1e-3) / 1e12
    Cf32 = C[ridx].cpu().numpy().astype(np.float64)

    # errors against FP64 on sampled rows
    C64 = A[ridx].cpu().numpy().astype(np.float64) @ B.cpu().numpy().astype(np.float64)
    e3, ef = np.abs(Cx3 - C64), np.abs(Cf32 - C64)
    err = {"rows_sampled": int(args.err_rows),
           "bf16x3_max_abs": float(e3.max()), "bf16x3_rms": float(np.sqrt((e3 ** 2).mean())),
           "fp32_fma_max_abs": float(ef.max()), "fp32_fma_rms": float(np.sqrt((ef ** 2).mean())),
           "ratio_max": float(e3.max() / ef.max()),
           "ratio_rms": float(np.sqrt((e3 ** 2).mean()) / np.sqrt((ef ** 2).mean()))}
    e6 = np.abs(Cx6 - C64)
    err["bf16x6_max_abs"] = float(e6.max())
    err["bf16x6_rms"] = float(np.sqrt((e6 ** 2).mean()))
    err["bf16x6_ratio_max"] = float(e6.max() / ef.max())
    err["bf16x6_ratio_rms"] = float(np.sqrt((e6 ** 2).mean()) / np.sqrt((ef ** 2).mean()))

    mfma_flops = 3 * flops           # three bf16 products per FP32 multiply-add
    achieved = mfma_flops / (core_ms * 1e-3) / 1e12
    line = base_line(args, round(tflops_x3, 2), "TFLOP/s", round(ms_x3, 4),
                     "f32 (split-bf16x3 MFMA, f32 accumulate)",
                     {"workload": "sgemm %dx%dx%d row-major f32, split-bf16x3 vs plain fp32-fma" % (n, n, n),
                      "M": n, "N": n, "K": n})
    line["roofline"] = {"bound": "mfma", "achieved": round(achieved, 1), "peak": PEAK_BF16_TFLOPS,
                        "unit": "TFLOP/s", "frac": round(achieved / PEAK_BF16_TFLOPS, 4),
                        "traffic": read_traffic(), "kernel": "gemm_x3_kernel", "kernel_ms": round(core_ms, 4),
                        "algorithmic_flops_per_launch": mfma_flops}
    line["fp32_fma"] = {"tflops": round(tflops_f32, 2), "ms": round(ms_f32, 3),
                        "frac_of_fp32_peak": round(tflops_f32 / PEAK_FP32_TFLOPS, 4)}
    line["speedup_vs_fp32"] = round(tflops_x3 / tflops_f32, 3)
    line["bf16x6"] = {"tflops": round(tflops_x6, 2), "ms": round(ms_x6, 3),
                      "speedup_vs_fp32": round(tflops_x6 / tflops_f32, 3),
                      "split3_prepass_ms": round(prep6_ms, 4),
                      "roofline": {"bound": "mfma", "achieved": round(6 * flops / (core6_ms * 1e-3) / 1e12, 1),
                                   "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                                   "frac": round(6 * flops / (core6_ms * 1e-3) / 1e12 / PEAK_BF16_TFLOPS, 4),
                                   "kernel": "gemm_x6_kernel", "kernel_ms": round(core6_ms, 4)},
                      "tile_variant": {"exhaustive": v6, "params": chai.gemm_variants(chai.GEMM_BF16X6)[v6]},
                      "note": "three bf16 pieces per operand, six MFMA products in one pass"}
    line["auto_family"] = "bf16x3" if n >= 16384 else "bf16x6"
    # the target: a split-BF16 form at >= 2x the FP32-FMA kernel with <= 2x its
    # FP64-referenced error, both in max and rms over the sampled rows
    line["fp32_level_target"] = {
        name: {"speedup_vs_fp32": round(sp, 3), "err_ratio_max": round(rmx, 3), "err_ratio_rms": round(rrms, 3),
               "met": bool(sp >= 2.0 and rmx <= 2.0 and rrms <= 2.0)}
        for name, sp, rmx, rrms in (("bf16x3", tflops_x3 / tflops_f32, err["ratio_max"], err["ratio_rms"]),
                                    ("bf16x6", tflops_x6 / tflops_f32, err["bf16x6_ratio_max"],
                                     err["bf16x6_ratio_rms"]))}
    line["split_prepass_ms"] = round(prep_ms, 4)
    line["errors_vs_fp64"] = err
    line["autotune_cold_start_s"] = cold
    line["tile_variant"] = {"exhaustive": cold["exhaustive_variant"],
                            "params": chai.gemm_variants()[cold["exhaustive_variant"]]}
    line["wall_s_timed"] = round(wall, 3)
    if not args.no_cpu:
        ref = cpu_reference_gemm(768)
        line["cpu_baseline"] = ref if ref is not None else cpu_baseline_gemm(args.cpu_size)
        if ref is not None:
            line["cpu_baseline_port"] = cpu_baseline_gemm(args.cpu_size)
    if not args.no_legs:
        del A, B, C
        line["legs"] = run_legs(args)
    emit(line)


LEG_DROP = ("metric", "n_gpus", "steps", "warmup", "higher_is_better", "scaling", "vs_baseline", "data")


def run_legs(args):
    """The other configurations measured in the same run, each with its own
    roofline and CPU baseline: FP64 and mixed i32 x f32 GEMM 4096^3, the
    4096-instance ragged batch, the fused MC trace with the Philox stream and
    GEMV 16384^2, the fused elementwise stream and the gather stencil."""
    import copy
    import gc
    import torch
    la = copy.copy(args)
    la.steps, la.warmup, la.cpu_budget = min(args.steps, 10), min(args.warmup, 2), 5.0
    legs = {}
    for name, fn in (("dgemm_f64_4096", lambda: run_dense(la, False, leg=True)),
                     ("mixed_i32xf32_4096", lambda: run_dense(la, True, leg=True)),
                     ("ragged_4096", lambda: run_ragged(la, leg=True)),
                     ("mc_philox_gemv", lambda: run_mc(la, leg=True)),
                     ("stream_trace_2e28", lambda: run_stream(la, leg=True)),
                     ("gather_stencil_8192", lambda: run_gather(la, leg=True))):
        t0 = time.perf_counter()
        d = fn()
        legs[name] = {k: v for k, v in d.items() if k not in LEG_DROP}
        legs[name]["steps"] = la.steps
        legs[name]["wall_s"] = round(time.perf_counter() - t0, 1)
        print("# leg %s: %.1f s" % (name, time.perf_counter() - t0), file=sys.stderr, flush=True)
        gc.collect()
        torch.cuda.empty_cache()
    return legs


# ------------------------------------------------------------------ ragged
def run_ragged(args, leg=False):
    import numpy as np
    import torch
    import chai_amd as chai
    from chai_amd import ragged
    from chai_amd.workloads import RAGGED_EPILOGUE, ragged_sizes
    chai.init(0)
    sizes = ragged_sizes()
    batch = []
    for i, k in enumerate(sizes):
        A, B = uniform((k, k), 42 + 2 * i), uniform((k, k), 43 + 2 * i)
        batch.append((A, B, torch.empty(k, k, device="cuda")))
    epi = chai.epilogue(*RAGGED_EPILOGUE)
    items = [ragged.item(*x) for x in batch]
    t0 = time.perf_counter()
    v, _ = ragged.tune(items, epi=epi, trials=2)
    tune_s = time.perf_counter() - t0
    plan = ragged.Plan(items, epi=epi, variant=v)
    sp, co = [], []
    for _ in range(1 + args.warmup + 3):           # per-launch hipEvents of eager runs, after a warm-up
        plan.run()
        a, b = plan.last_timing()
        sp.append(a); co.append(b)
    split_ms, core_ms = sum(sp[-3:]) / 3, sum(co[-3:]) / 3
    info = plan.info()
    ms = timed(lambda: plan.run(graph=True), args.steps, args.warmup)
    flops = info["useful_flops"]
    # per-instance launches of the same kernels (fused epilogue), as the reference runs them
    dv = v if v < len(chai.gemm_variants()) else 2     # the f32-fed ragged tiles have no dense twin
    ms_pi = timed(lambda: [chai.gemm(A, B, C, mode=chai.GEMM_BF16X3, epi=epi, variant=dv) for A, B, C in batch],
                  1, 1)
    line = base_line(args, round(flops / (ms * 1e-3) / 1e12, 2), "TFLOP/s", round(ms, 4),
                     "f32 (split-bf16x3 MFMA, f32 accumulate)",
                     {"workload": "ragged batch: 4096 square f32 GEMMs, n log-uniform in [32,1024] step 16 "
                                  "(seed 42), epilogue max(C*0.5+1,0), one persistent launch in a hipGraph",
                      "instances": len(sizes), "sum_2n3": flops})
    line["roofline"] = {"bound": "mfma", "achieved": round(3 * info["padded_flops"] / (core_ms * 1e-3) / 1e12, 1),
                        "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                        "frac": round(3 * info["padded_flops"] / (core_ms * 1e-3) / 1e12 / PEAK_BF16_TFLOPS, 4),
                        "traffic": None, "kernel": "ragged_x3_kernel", "kernel_ms": round(core_ms, 4)}
    line["split_prepass_ms"] = round(split_ms, 4)
    line["tile_variant"] = v
    line["tile_variant_kind"] = "f32-fed tiles (no split pre-pass)" if info["split_blocks"] == 0 else \
        "split pre-pass + panel tiles"
    line["tune_s"] = round(tune_s, 2)
    line["plan"] = info
    line["per_instance_ms"] = round(ms_pi, 3)
    line["speedup_vs_per_instance"] = round(ms_pi / ms, 2)
    if not args.no_cpu:
        from oracle import oracle as o
        rng = np.random.default_rng(1)
        sub = rng.choice(len(sizes), 64, replace=False)
        t0 = time.perf_counter()
        f_sub = 0.0
        for i in sub:
            k = int(sizes[i])
            Ah = batch[i][0].cpu().numpy()
            Bh = batch[i][1].cpu().numpy()
            Ch = o.matmul_interp(Ah, Bh)
            np.maximum(Ch.astype(np.float32) * np.float32(0.5) + np.float32(1), 0)
            f_sub += 2.0 * k ** 3
            if time.perf_counter() - t0 > args.cpu_budget:
                break
        dt = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": f_sub / dt / 1e12, "unit": "TFLOP/s", "cores": 1, "kind": "port",
                                "sample": "the oracle's interpreter-semantics matmul + epilogue on a random "
                                          "subsample of the instances (%.2e flop, %.1f s, 1 core)" % (f_sub, dt)}
        line["cpu_baseline"].update(host_info())
    if leg:
        return line
    emit(line)


# ----------------------------------------------------------------- autotune
def run_autotune(args):
    """The learned-autotune workload: the exhaustive timed search of every
    tile variant of the four modelled families (BF16X3, FMA, F64, GEMV) on
    the training shapes, one MLP trained on all of them, then per family on
    the held-out shapes: the model's choice against the search's best
    (throughput regret) and the wall time of one inference against the
    search's (cold start)."""
    import numpy as np
    import torch
    import chai_amd as chai
    from chai_amd import tuner
    chai.init(0)
    ws = tuner.Workspace()
    train = tuner.random_shapes(args.train_shapes, seed=2024)
    test = tuner.random_shapes(args.test_shapes, seed=2025)
    fams = tuner.FAMILIES
    t_train, t_test, cold_ex = {}, {}, {}
    t0 = time.perf_counter()
    for f in fams:
        name = tuner.FAMILY_NAMES[f]
        t_train[f] = []
        for i, s in enumerate(train):
            t_train[f].append(tuner.time_variants(ws, s, family=f)[0])
            if i % 250 == 0:
                print("# %s: timed %d/%d train shapes (%.0f s)" % (name, i, len(train), time.perf_counter() - t0),
                      file=sys.stderr, flush=True)
        t_test[f], cold_ex[f] = [], []
        for s in test:
            ms, secs = tuner.time_variants(ws, s, family=f)
            t_test[f].append(ms)
            cold_ex[f].append(secs)
        t_train[f], t_test[f] = np.array(t_train[f]), np.array(t_test[f])
    collect_s = time.perf_counter() - t0
    Xs, Ys = [], []
    for f in fams:
        X, Y = tuner.dataset(train, t_train[f], f)
        Xs.append(X)
        Ys.append(Y)
    X, Y = np.concatenate(Xs), np.concatenate(Ys)
    mu, sd = float(Y.mean()), float(Y.std())
    m = tuner.MLP(tuner.NFEAT, seed=0)
    t1 = time.perf_counter()
    m.fit(torch.from_numpy(X).cuda(), torch.from_numpy((Y - mu) / sd).cuda(), epochs=args.steps * 20,
          batch=256, lr=2e-3)
    train_s = time.perf_counter() - t1
    if args.save_model:
        tuner.save_model(m, args.save_model, mu, sd)
    per = {}
    cold_all_m, cold_all_ex = [], []
    for f in fams:
        tuner.predict(m, test[0], f)          # one-time setup
        choice, cold_m = [], []
        for s in test:
            v, _, secs = tuner.predict(m, s, f)
            choice.append(v)
            cold_m.append(secs)
        tt = t_test[f]
        reg = tuner.regret(choice, tt)
        fixed = int(np.argmin(np.log(t_train[f]).sum(0)))
        per[tuner.FAMILY_NAMES[f]] = {
            "variants": int(tt.shape[1]),
            "regret_mean": float(reg.mean()), "regret_median": float(np.median(reg)),
            "regret_p90": float(np.quantile(reg, 0.9)),
            "perf_vs_exhaustive_mean": float(np.mean(tt.min(1) / tt[np.arange(len(test)), choice])),
            "regret_fixed_best_variant": float(tuner.regret([fixed] * len(test), tt).mean()),
            "fixed_variant": fixed,
            "regret_random_variant": float(np.mean([tuner.regret([v] * len(test), tt).mean()
                                                    for v in range(tt.shape[1])])),
            "cold_start_model_s_mean": float(np.mean(cold_m)),
            "cold_start_exhaustive_s_mean": float(np.mean(cold_ex[f])),
            "cold_start_speedup": float(np.mean(cold_ex[f]) / np.mean(cold_m))}
        cold_all_m += cold_m
        cold_all_ex += cold_ex[f]
    cm = float(np.mean(cold_all_m))
    line = base_line(args, round(cm, 6), "s", round(cm * 1e3, 4), "f32 (MLP 3x64)",
                     {"workload": "learned autotune: %d train / %d held-out shapes, M,N,K log-uniform [64,16384] "
                                  "step 16; BF16X3 / FMA / F64 / GEMV tile variants" % (len(train), len(test))},
                     higher=False)
    line["metric"] = METRIC
    line["autotune"] = {
        "families": per,
        "regret_mean_worst_family": max(v["regret_mean"] for v in per.values()),
        "cold_start_model_s_mean": cm, "cold_start_exhaustive_s_mean": float(np.mean(cold_all_ex)),
        "cold_start_speedup": float(np.mean(cold_all_ex) / cm),
        "train_rows": int(X.shape[0]), "train_s": round(train_s, 2), "collect_s": round(collect_s, 1)}
    emit(line)


# ----------------------------------------------------------------------- mc
def run_mc(args, leg=False):
    import torch
    import chai_amd as chai
    from chai_amd.workloads import mc_trace
    chai.init(0)
    n = 1 << 24
    t, m = mc_trace(n)
    prog = t.compile()
    out = torch.zeros(1, device="cuda")
    ms_mc = timed(lambda: prog.run({m: out}), args.steps, args.warmup)
    z = torch.empty(n, device="cuda")
    ms_norm = timed(lambda: chai.rng_fill(z, seed=7, normal=True), args.steps, args.warmup)
    N = 16384
    A, x = uniform((N, N), 42), uniform((N,), 42 + N * N)
    y = torch.empty(N, device="cuda")
    # the GEMV family's tile search for this shape (chai.gemv then uses its choice)
    gv, _ = chai.gemm_tune(A, x.view(-1, 1), y.view(-1, 1), mode=chai.GEMM_GEMV, trials=3)
    ms_gemv = timed(lambda: chai.gemv(A, x, y=y), args.steps, args.warmup)
    gemv_bytes = 4.0 * (N * N + 2 * N)
    gbs = gemv_bytes / (ms_gemv * 1e-3) / 1e9
    line = base_line(args, round(n / (ms_mc * 1e-3) / 1e9, 3), "Gelem/s", round(ms_mc, 4), "f32",
                     {"workload": "fused MC trace 2^24 Philox4x32-10 normals (seed 7) -> payoff mean; "
                                  "f32 GEMV 16384^2"})
    line["mc_mean"] = float(out.item())
    line["philox_normals_2^24_ms"] = round(ms_norm, 4)
    line["philox_normals_GBs"] = round(4.0 * n / (ms_norm * 1e-3) / 1e9, 1)
    # steady state of the raw stream: 2^28 f32 (1 GiB written)
    big = torch.empty(1 << 28, device="cuda")
    ms_u = timed(lambda: chai.rng_fill(big, seed=7), args.steps, args.warmup)
    ms_n = timed(lambda: chai.rng_fill(big, seed=7, normal=True), args.steps, args.warmup)
    line["philox_uniform_2^28"] = {"ms": round(ms_u, 4), "GBs": round(4.0 * 2 ** 28 / (ms_u * 1e-3) / 1e9, 1),
                                   "frac_hbm": round(4.0 * 2 ** 28 / (ms_u * 1e-3) / 1e9 / PEAK_HBM_GBS, 3)}
    line["philox_normals_2^28"] = {"ms": round(ms_n, 4), "GBs": round(4.0 * 2 ** 28 / (ms_n * 1e-3) / 1e9, 1),
                                   "frac_hbm": round(4.0 * 2 ** 28 / (ms_n * 1e-3) / 1e9 / PEAK_HBM_GBS, 3)}
    # the same 1 GiB written by torch's fill kernel: the write bandwidth this part reaches
    ms_fill = timed(lambda: big.fill_(1.0), args.steps, args.warmup)
    line["torch_fill_2^28_GBs"] = round(4.0 * 2 ** 28 / (ms_fill * 1e-3) / 1e9, 1)
    del big
    line["gemv_ms"] = round(ms_gemv, 4)
    line["gemv_variant"] = {"tuned": gv, "params": chai.gemm_variants(chai.GEMM_GEMV)[gv][:2]}
    line["roofline"] = {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": round(gbs / PEAK_HBM_GBS, 4), "traffic": None, "kernel": "gemv_rows"}
    if not args.no_cpu:
        from oracle import oracle as o
        k = 1 << 20
        tc, mc = mc_trace(k)
        t0 = time.perf_counter()
        o.trace_eval(tc, {})
        dt = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": k / dt / 1e9, "unit": "Gelem/s", "cores": 1, "kind": "port",
                                "sample": "oracle trace evaluation of the MC trace on 2^20 elements (%.1f s)" % dt}
        line["cpu_baseline"].update(host_info())
    if leg:
        return line
    emit(line)


# ------------------------------------------------------------------ gather
def run_gather(args, leg=False):
    """sample/filter.cpp at N = 8192: one fused kernel of five gather2_floor
    reads; algorithmic bytes 4 read + 4 written per element (the stencil's
    neighbours come from L2 / the same rows)"""
    import numpy as np
    import torch
    import chai_amd as chai
    from chai_amd.trace import set_jit
    from chai_amd.workloads import filter_expected, filter_trace
    chai.init(0)
    N = 8192
    t, A, B = filter_trace(N)
    a = uniform((N, N), 42)
    b = torch.empty(N * N, device="cuda")
    res = {}
    for eng in ("jit", "vm"):
        set_jit(eng == "jit")
        prog = t.compile()
        res[eng] = timed(lambda: prog.run({A: a.view(-1), B: b}), args.steps, args.warmup)
    set_jit(True)
    ms = res["jit"]
    gbs = 8.0 * N * N / (ms * 1e-3) / 1e9
    # correctness of the measured kernel on a corner block, against the periodic stencil
    an = a.cpu().numpy()
    ok = bool(np.array_equal(b.view(N, N).cpu().numpy()[:64, :64],
                             filter_expected(an[None])[0][:64, :64]))
    at = a.view(N, N)
    ms_torch = timed(lambda: at + torch.roll(at, 1, 1) + torch.roll(at, -1, 1) + torch.roll(at, 1, 0)
                     + torch.roll(at, -1, 0), args.steps, args.warmup)
    line = base_line(args, round(gbs, 1), "GB/s", round(ms, 4), "f32",
                     {"workload": "gather2_floor 5-point stencil (sample/filter.cpp) on 8192x8192 f32"})
    line["metric"] = "fused gather stencil GB/s (algorithmic bytes)"
    line["vm_ms"] = round(res["vm"], 4)
    line["torch_roll_stencil_ms"] = round(ms_torch, 4)
    line["matches_periodic_stencil"] = ok
    line["roofline"] = {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": round(gbs / PEAK_HBM_GBS, 4), "traffic": read_traffic(PMC_SUMMARY_GATHER),
                        "kernel": "chai_trace_kernel"}
    if not args.no_cpu:
        from oracle import oracle as o
        n = 2048
        tc, Ac, Bc = filter_trace(n)
        x = np.random.default_rng(0).standard_normal((1, n, n)).astype(np.float32)
        t0 = time.perf_counter()
        reps = 0
        while time.perf_counter() - t0 < min(10.0, args.cpu_budget):
            o.trace_eval(tc, {Ac.index: x})
            reps += 1
        dt = (time.perf_counter() - t0) / reps
        line["cpu_baseline"] = {"value": round(8.0 * n * n / dt / 1e9, 3), "unit": "GB/s", "cores": 1,
                                "kind": "port", "sample": "oracle trace evaluation of the stencil on 2048x2048, "
                                                          "%d runs, %.2f s each" % (reps, dt)}
        line["cpu_baseline"].update(host_info())
    if leg:
        return line
    emit(line)


# ------------------------------------------------------------ dgemm / mixed
def run_dense(args, mixed, leg=False):
    import numpy as np
    import torch
    import chai_amd as chai
    chai.init(0)
    n = 4096
    if mixed:
        W = torch.empty(n, n, dtype=torch.int32, device="cuda")
        chai.rng_fill(W, seed=42, prec=chai.U32)               # raw Philox words
        A = (W & 255) - 128                                    # lo + (w mod 256) in [-128, 127]
        B = uniform((n, n), 42 + n * n)
        fn = lambda: chai.gemm(A, B, C=C, mode=chai.GEMM_BF16X3)   # noqa: E731
        C = torch.empty(n, n, device="cuda")
        dtype, peak, bound = "i32 x f32 -> f32 (bf16x3 MFMA)", PEAK_BF16_TFLOPS, "mfma"
    else:
        A = uniform((n, n), 42, torch.float64)
        B = uniform((n, n), 42 + n * n, torch.float64)
        C = torch.empty(n, n, dtype=torch.float64, device="cuda")
        fn = lambda: chai.gemm(A, B, C=C, mode=chai.GEMM_F64)     # noqa: E731
        dtype, peak, bound = "f64 (FP64 MFMA)", PEAK_FP64_TFLOPS, "mfma"
    # tile search first (the reference tunes every new exogenous key); the timed
    # steps run the variant it picks
    mode = chai.GEMM_BF16X3 if mixed else chai.GEMM_F64
    tv, _ = chai.gemm_tune(A, B, C, mode=mode, trials=2)
    # core kernel durations recorded by the library over the timed steps
    chai.profile(True)
    ms = timed(fn, args.steps, args.warmup)
    core_ms = chai.gemm_profile_mean(min(args.steps, 64))[1]
    chai.profile(False)
    tf = 2.0 * n ** 3 / (ms * 1e-3) / 1e12
    # MFMA work per f32 multiply-add: the ints in [-128, 127] are exact in bf16,
    # so their lo panel is zero and the tiles run two products (ah*bl, ah*bh)
    products = 2 if mixed else 1
    ach = products * 2.0 * n ** 3 / (core_ms * 1e-3) / 1e12
    rows = np.random.default_rng(0).choice(n, 8, replace=False)
    ridx = torch.from_numpy(rows).cuda()
    ref = A[ridx].cpu().numpy().astype(np.float64) @ B.cpu().numpy().astype(np.float64)
    err = float(np.abs(C[ridx].cpu().numpy().astype(np.float64) - ref).max())
    line = base_line(args, round(tf, 2), "TFLOP/s", round(ms, 4), dtype,
                     {"workload": ("mixed i32[-128,127] x f32 " if mixed else "dgemm f64 ") + "4096^3"})
    line["roofline"] = {"bound": bound, "achieved": round(ach, 1), "peak": peak, "unit": "TFLOP/s",
                        "frac": round(ach / peak, 4), "traffic": None, "kernel_ms": round(core_ms, 4),
                        "mfma_products_per_mad": products}
    line["tile_variant"] = {"tuned": tv, "params": chai.gemm_variants(mode)[tv]}
    line["max_abs_err_vs_fp64"] = err
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline_dense(mixed, args.cpu_budget)
    if leg:
        return line
    emit(line)


def cpu_baseline_dense(mixed, budget):
    """The oracle's interpreter-semantics product on one host core for the
    dense legs: f64 (DGEMM) or i32 x f32 (mixed) at 384^3"""
    import numpy as np
    from oracle import oracle as o
    n = 384
    B = o.rng_fill(1230, 4, 4, 43, n, n, limits=(-1.0, 1.0)).reshape(n, n)
    if mixed:
        A = (np.random.default_rng(0).integers(-128, 128, (n, n))).astype(np.int32)
    else:
        A = o.rng_fill(1230, 8, 2, 42, n, n, limits=(-1.0, 1.0)).reshape(n, n)
        B = B.astype(np.float64)
    t0 = time.perf_counter()
    reps = 0
    while reps == 0 or time.perf_counter() - t0 < min(3.0, budget):
        o.matmul_interp(A, B)
        reps += 1
    dt = (time.perf_counter() - t0) / reps
    out = {"value": 2.0 * n ** 3 / dt / 1e12, "unit": "TFLOP/s", "cores": 1, "kind": "port",
           "sample": "the oracle's interpreter-semantics %s matmul %d^3, 1 core, %.2f s"
                     % ("i32 x f32" if mixed else "f64", n, dt)}
    out.update(host_info())
    return out


# ------------------------------------------------------------------ errors per shape
def run_errors(args):
    """Split-BF16x3 and plain FP32-FMA GEMM errors against the FP64 product, per shape.
    M = N = 1024 (every element checked) and K from 256 to 16384, for operands uniform in
    [-1, 1) (config inputs, seed 42) and in [0, 1) (no cancellation).  Per shape: max and rms
    absolute error, the max of |C - C64| / sum_k |a b| (the quantity the written bounds cap),
    and the bf16x3 / fp32 ratios."""
    import numpy as np
    import torch
    import chai_amd as chai
    chai.init(0)
    m = n = 1024
    rows = []
    for lo in (-1.0, 0.0):
        for k in (256, 1024, 2048, 4096, 8192, 16384):
            A = torch.empty(m, k, device="cuda")
            B = torch.empty(k, n, device="cuda")
            chai.rng_fill(A, seed=42, limits=(lo, 1.0), W=k, H=m)
            chai.rng_fill(B, seed=42 + m * k, limits=(lo, 1.0), W=n, H=k)
            A64, B64 = A.cpu().numpy().astype(np.float64), B.cpu().numpy().astype(np.float64)
            C64 = A64 @ B64
            scale = np.abs(A64) @ np.abs(B64)
            r = {"M": m, "N": n, "K": k, "inputs": "uniform[%g,1)" % lo}
            for name, mode in (("bf16x3", chai.GEMM_BF16X3), ("bf16x6", chai.GEMM_BF16X6),
                               ("fp32_fma", chai.GEMM_FMA)):
                e = np.abs(chai.gemm(A, B, mode=mode).cpu().numpy().astype(np.float64) - C64)
                r[name] = {"max_abs": float(e.max()), "rms": float(np.sqrt((e ** 2).mean())),
                           "max_rel_to_sum_abs": float((e / scale).max())}
            r["ratio_max"] = r["bf16x3"]["max_abs"] / r["fp32_fma"]["max_abs"]
            r["ratio_rms"] = r["bf16x3"]["rms"] / r["fp32_fma"]["rms"]
            r["x6_ratio_max"] = r["bf16x6"]["max_abs"] / r["fp32_fma"]["max_abs"]
            r["x6_ratio_rms"] = r["bf16x6"]["rms"] / r["fp32_fma"]["rms"]
            rows.append(r)
            print("K=%5d %-16s x3/fp32 max %.2f rms %.2f | x6/fp32 max %.2f rms %.2f"
                  % (k, r["inputs"], r["ratio_max"], r["ratio_rms"], r["x6_ratio_max"], r["x6_ratio_rms"]),
                  flush=True)
    print(json.dumps({"errors_vs_fp64_per_shape": rows}))


# ------------------------------------------------------------------ stream
def run_stream(args, leg=False):
    """An HBM-bound fused elementwise trace, C = max(A * 0.5f + B, 0) over 2^28 f32
    (1 GiB per array): 8 B read + 4 B written per element, one fused kernel.  Beside it:
    torch's two-input elementwise kernel and a plain device copy (8 B per element),
    the bandwidths this part reaches for the same access pattern."""
    import numpy as np
    import torch
    import chai_amd as chai
    from chai_amd import trace as tr
    chai.init(0)
    n = 1 << 28
    t, A, B, C = stream_trace(n)
    a, b = uniform((n,), 42), uniform((n,), 43)
    c = torch.empty(n, device="cuda")
    res = {}
    for eng in ("jit", "vm"):
        tr.set_jit(eng == "jit")
        prog = t.compile()
        res[eng] = timed(lambda: prog.run({A: a, B: b, C: c}), args.steps, args.warmup)
    tr.set_jit(True)
    ms = res["jit"]
    k = 1 << 20
    an, bn = a[:k].cpu().numpy(), b[:k].cpu().numpy()
    ok = bool(np.array_equal(c[:k].cpu().numpy(), np.maximum(an * np.float32(0.5) + bn, np.float32(0))))
    gbs = 12.0 * n / (ms * 1e-3) / 1e9
    ms_torch = timed(lambda: torch.add(b, a, alpha=0.5, out=c), args.steps, args.warmup)
    ms_copy = timed(lambda: c.copy_(a), args.steps, args.warmup)
    line = base_line(args, round(gbs, 1), "GB/s", round(ms, 4), "f32",
                     {"workload": "fused elementwise trace C = max(A*0.5f + B, 0) over 2^28 f32"})
    line["metric"] = "fused elementwise trace GB/s (algorithmic bytes)"
    line["vm_ms"] = round(res["vm"], 4)
    line["matches_numpy_first_2^20"] = ok
    line["torch_add_GBs"] = round(12.0 * n / (ms_torch * 1e-3) / 1e9, 1)
    line["device_copy_GBs"] = round(8.0 * n / (ms_copy * 1e-3) / 1e9, 1)
    line["roofline"] = {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": round(gbs / PEAK_HBM_GBS, 4), "traffic": read_traffic(PMC_SUMMARY_STREAM),
                        "kernel": "chai_trace_kernel"}
    # a fused reduction trace over the same arrays: s = sum(A * B) (8 B read per
    # element, FP64 block partials combined in a fixed order), beside the
    # library's plain sum and torch's
    tr_ = tr.Trace()
    Ar, Br = tr_.input(chai.F32, n), tr_.input(chai.F32, n)
    Sr = tr_.output(chai.F32, 1)
    tr_.assign(Sr, tr.sum_(Ar * Br))
    prog_r = tr_.compile()
    s_out = torch.empty(1, device="cuda")
    ms_r = timed(lambda: prog_r.run({Ar: a, Br: b, Sr: s_out}), args.steps, args.warmup)
    ref = float(torch.dot(a.double(), b.double()))
    ms_sum = timed(lambda: chai.reduce(a, "sum"), args.steps, args.warmup)
    ms_tsum = timed(lambda: torch.sum(a), args.steps, args.warmup)
    line["reduction_trace_2^28"] = {
        "trace": "s = sum(A * B)", "ms": round(ms_r, 4), "GBs": round(8.0 * n / (ms_r * 1e-3) / 1e9, 1),
        "frac_hbm": round(8.0 * n / (ms_r * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
        "rel_err_vs_fp64": abs(float(s_out.item()) - ref) / abs(ref),
        "chai_reduce_sum_GBs": round(4.0 * n / (ms_sum * 1e-3) / 1e9, 1),
        "torch_sum_GBs": round(4.0 * n / (ms_tsum * 1e-3) / 1e9, 1)}
    if not args.no_cpu:
        from oracle import oracle as o
        kc = 1 << 22
        tc, Ac, Bc, Cc = stream_trace(kc)
        xa = np.random.default_rng(0).standard_normal(kc).astype(np.float32)
        t0 = time.perf_counter()
        reps = 0
        while time.perf_counter() - t0 < min(5.0, args.cpu_budget):
            o.trace_eval(tc, {Ac.index: xa, Bc.index: xa})
            reps += 1
        dt = (time.perf_counter() - t0) / reps
        line["cpu_baseline"] = {"value": round(12.0 * kc / dt / 1e9, 3), "unit": "GB/s", "cores": 1,
                                "kind": "port", "sample": "oracle trace evaluation on 2^22 elements, "
                                                          "%d runs, %.2f s each" % (reps, dt)}
        line["cpu_baseline"].update(host_info())
    if leg:
        return line
    emit(line)


def main():
    args = parse()
    if args.pmc_summary:
        pmc_summary(*args.pmc_summary, workload=args.workload)
        return
    if args.kernel_stats:
        kernel_stats(*args.kernel_stats)
        return
    if args.pmc_table:
        pmc_table(args.pmc_table)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: --gpus %d is not supported by this build "
              "(the GEMM would shard by 2-D output tiles across GPUs, all-gathering B panels over xGMI)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.profile_run and args.workload == "ragged":
        profile_run_ragged(args)
    elif args.profile_run and args.workload in ("gather", "stream"):
        profile_run_gather(args)
    elif args.profile_run:
        profile_run(args)
    elif args.workload == "sgemm":
        run_sgemm(args)
    elif args.workload == "ragged":
        run_ragged(args)
    elif args.workload == "autotune":
        run_autotune(args)
    elif args.workload == "mc":
        run_mc(args)
    elif args.workload == "stream":
        run_stream(args)
    elif args.workload == "errors":
        run_errors(args)
    elif args.workload == "gather":
        run_gather(args)
    else:
        run_dense(args, args.workload == "mixed")


if __name__ == "__main__":
    main()
