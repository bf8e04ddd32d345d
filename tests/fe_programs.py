"""Inputs and runners for the reference-API test programs (tests/frontend/).

fe_cases (tests/frontend/fe_cases.cpp) and the reference's own sample
programs are built against the reference's library compiled here
(oracle/frontend.mk) and the MI355X device (chai_amd/device/).  Each run
happens in a fresh process: the reference's scheduler is a process-wide
singleton that is initialised once per program.
"""
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "fe_bin")
FE_CASES = os.path.join(BIN, "fe_cases")
ORACLE_BACKEND = os.path.join(BIN, "liboracle_backend.so")
REAL_LIB = os.path.join(ROOT, "chai_amd", "libchai_mi355x.so")

# the reference's sample programs built by tests/frontend/Makefile
SAMPLES = ["md5", "sum", "sum_pth", "loopsum", "matmul", "matmul_pth", "filter", "index", "sum_vec", "matmul_vec",
           "matmul64", "cg", "kirch", "mingle", "pi", "monte"]
# samples that draw random arrays: the interpreter's generator is glibc
# random_r seeded with the seed XOR the generator object's address
# (interp/InterpRNG.cpp:11-23), so its values change from run to run; the
# device draws Chai's Random123 stream (the JIT's), which is deterministic
RNG_SAMPLES = ["pi", "monte"]


def built():
    return os.path.exists(FE_CASES) and all(os.path.exists(os.path.join(BIN, "sample_" + s)) for s in SAMPLES)


def _u(g, n, lo=-1.0, hi=1.0):
    return g.uniform(lo, hi, n).astype(np.float32)


def case_inputs(case, seed=7):
    """{name: array} written as <case>.<name>.bin; shapes as fe_cases.cpp reads them."""
    g = np.random.default_rng(seed)
    if case == "mm_f32":
        # config 1: 256^3, A and B uniform in [-1, 1) from Chai's Philox stream, seed 42
        from oracle import oracle as o
        A = o.rng_fill(1230, 4, 4, 42, 256, 256, limits=(-1.0, 1.0))
        B = o.rng_fill(1230, 4, 4, 42 + 65536, 256, 256, limits=(-1.0, 1.0))
        return {"A": A.ravel(), "B": B.ravel()}
    if case == "mm_shapes":
        M, K, N = 96, 160, 72
        return {"A": _u(g, M * K), "At": _u(g, K * M), "B": _u(g, K * N), "Bt": _u(g, N * K),
                "C0": _u(g, M * N), "x": _u(g, K)}
    if case == "mm_int":
        M, K, N = 64, 48, 40
        ua = g.integers(0, 2**32, M * K, dtype=np.uint64).astype(np.uint32)
        ub = g.integers(0, 2**32, K * N, dtype=np.uint64).astype(np.uint32)
        ia = g.integers(-2**31, 2**31, M * K, dtype=np.int64).astype(np.int32)
        ib = g.integers(-2**31, 2**31, K * N, dtype=np.int64).astype(np.int32)
        # the first rows are small, so some sums stay inside int32
        ia[:8 * K] = g.integers(-300, 300, 8 * K).astype(np.int32)
        ib[:] = np.where(np.arange(K * N) % 3 == 0, ib, g.integers(-300, 300, K * N)).astype(np.int32)
        return {"UA": ua, "UB": ub, "IA": ia, "IB": ib, "FB": _u(g, K * N)}
    if case == "mm_f64":
        return {"A": g.uniform(-1, 1, 64 * 80), "B": g.uniform(-1, 1, 80 * 48)}
    if case == "ew":
        return {"A": _u(g, 64 * 48, -3, 3), "B": _u(g, 64 * 48, -3, 3)}
    if case == "red":
        n = 256 * 96
        return {"A": _u(g, n), "B": _u(g, n), "IA": g.integers(-2**20, 2**20, n).astype(np.int32)}
    if case == "ints":
        n = 128 * 16
        return {"UA": g.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32),
                "UB": g.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32),
                "IA": g.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32),
                "IB": g.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32)}
    if case == "gather":
        return {"A": _u(g, 64 * 32)}
    if case == "multi":
        return {"A": _u(g, 128 * 64), "B": _u(g, 128 * 64)}
    raise KeyError(case)


OUTPUTS = {
    "mm_f32": {"C": np.float32},
    "mm_shapes": {"C1": np.float32, "C2": np.float32, "C3": np.float32, "C4": np.float32, "y": np.float32},
    "mm_int": {"CUU": np.uint32, "CII": np.int32, "CUI": np.int32, "CIF": np.float32},
    "mm_f64": {"C": np.float64},
    "ew": {"E": np.float32, "F": np.float32, "G": np.float32},
    "red": {"r": np.float32, "ri": np.int32},
    "ints": {"X": np.uint32, "Y": np.int32},
    "gather": {"O": np.float32},
    "multi": {"D": np.float32, "E": np.float32, "s": np.float32},
}
CASES = list(OUTPUTS)


def env_for(device, lib=None):
    """Environment of a run: device None = interpreter only; 'oracle' = the
    MI355X device over the CPU oracle; 'missing' = the device with no
    library (every trace falls back); 'gpu' = the device over
    libchai_mi355x.so."""
    e = dict(os.environ)
    e["OCL_ICD_VENDORS"] = "/nonexistent"      # no OpenCL devices: interpreter + MI355X only
    e.pop("CHAI_MI355X", None)
    if device is None:
        return e
    e["CHAI_MI355X"] = "0"
    e["CHAI_MI355X_VERBOSE"] = "1"
    if device == "oracle":
        e["CHAI_MI355X_LIB"] = ORACLE_BACKEND
        e["CHAI_REAL_LIB"] = REAL_LIB
    elif device == "missing":
        e["CHAI_MI355X_LIB"] = os.path.join(BIN, "no_such_library.so")
    else:
        e["CHAI_MI355X_LIB"] = lib or REAL_LIB
    return e


def run_case(case, workdir, device=None, timeout=120):
    """Run fe_cases once; returns ({name: array}, stderr)."""
    os.makedirs(workdir, exist_ok=True)
    for name, a in case_inputs(case).items():
        np.ascontiguousarray(a).tofile(os.path.join(workdir, "%s.%s.bin" % (case, name)))
    r = subprocess.run([FE_CASES, case, workdir], env=env_for(device), capture_output=True, text=True,
                       timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError("fe_cases %s (%s) exited %d: %s" % (case, device, r.returncode, r.stderr[-2000:]))
    out = {name: np.fromfile(os.path.join(workdir, "%s.out.%s.bin" % (case, name)), dtype=dt)
           for name, dt in OUTPUTS[case].items()}
    return out, r.stderr


def run_sample(name, device=None, timeout=120):
    r = subprocess.run([os.path.join(BIN, "sample_" + name), "-i"], env=env_for(device), capture_output=True,
                       text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError("sample %s (%s) exited %d: %s" % (name, device, r.returncode, r.stderr[-2000:]))
    return r.stdout, r.stderr
