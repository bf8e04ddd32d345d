"""Learned tile choice for the GEMM and GEMV kernel families.

The reference chooses GEMM / GEMV kernel parameters by timed search only
(autotune/StandardEM.cpp:69-121, autotune/AutoTuneEM.cpp:48-131,
enqueue/EnqueueMatmulAuto.cpp:71-161).  Here the same search
(chai_hip_gemm_time_variant / chai_hip_gemm_tune) produces the training data of a
small MLP (chai_hip_mlp_*: 3 x 64 ReLU, FP32, trained by the library's HIP
kernels) that predicts every variant's core-kernel time from the shape, the
family and the variant's tile parameters; a new shape then gets its variant
from one inference (chai_hip_gemm_predict_variant) instead of a search, and
the search remains the fallback and measures the model's regret.  One model
covers five families: the split-BF16x3 MFMA tiles (8 variants), the
split-BF16x6 tiles (5), the FP32-FMA tiles (4), the FP64 MFMA tiles (3) and
the GEMV kernels (9: waves per row or row chunks x loads in flight).
"""
import ctypes
import os
import time

import numpy as np

from . import (F32, F64, GEMM_BF16X3, GEMM_BF16X6, GEMM_F64, GEMM_FMA, GEMM_GEMV, I32, GemmDesc,  # noqa: F401
               _stream, check, gemm_variants, init, lib)

# the families the model covers (the one-hot order of the features)
FAMILIES = (GEMM_BF16X3, GEMM_FMA, GEMM_F64, GEMM_GEMV, GEMM_BF16X6)
FAMILY_NAMES = {GEMM_BF16X3: "bf16x3", GEMM_FMA: "fma", GEMM_F64: "f64", GEMM_GEMV: "gemv", GEMM_BF16X6: "bf16x6"}
NFEAT = 25          # CHAI_GEMM_FEATURES
_vp, _sz = ctypes.c_void_p, ctypes.c_size_t


def _bind():
    L = lib()
    if getattr(L, "_mlp_bound", False):
        return L
    P = ctypes.POINTER
    L.chai_hip_mlp_create.argtypes = [ctypes.c_int, ctypes.c_uint64, P(_vp)]
    L.chai_hip_mlp_num_params.argtypes = [_vp, P(_sz)]
    L.chai_hip_mlp_get_params.argtypes = [_vp, _vp, _sz]
    L.chai_hip_mlp_set_params.argtypes = [_vp, _vp, _sz]
    L.chai_hip_mlp_forward.argtypes = [_vp, _vp, _sz, _vp, _vp]
    L.chai_hip_mlp_backward.argtypes = [_vp, _vp, _vp, _sz, _vp, _vp]
    L.chai_hip_mlp_update.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_float, _vp]
    L.chai_hip_mlp_train_step.argtypes = [_vp, _vp, _vp, _sz, ctypes.c_int, ctypes.c_float, _vp]
    L.chai_hip_mlp_release.argtypes = [_vp]
    L.chai_hip_gemm_features.argtypes = [P(GemmDesc), ctypes.c_int, _vp]
    L.chai_hip_gemm_predict_variant.argtypes = [P(GemmDesc), _vp, P(ctypes.c_int), _vp, _vp]
    L.chai_hip_gemm_time_variant.argtypes = [P(GemmDesc), _vp, _vp, _vp, ctypes.c_int, ctypes.c_int,
                                             P(ctypes.c_float), _vp]
    L._mlp_bound = True
    return L


class MLP:
    """Handle on a chai_hip_mlp (device parameters + Adam state)."""

    def __init__(self, inputs=NFEAT, seed=0):
        init()
        L = _bind()
        h = _vp()
        check(L.chai_hip_mlp_create(inputs, seed, ctypes.byref(h)), "chai_hip_mlp_create")
        self._h = h
        self.inputs = inputs
        n = _sz()
        check(L.chai_hip_mlp_num_params(h, ctypes.byref(n)), "chai_hip_mlp_num_params")
        self.num_params = n.value

    def params(self):
        p = np.empty(self.num_params, dtype=np.float32)
        check(_bind().chai_hip_mlp_get_params(self._h, p.ctypes.data, p.size), "chai_hip_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(_bind().chai_hip_mlp_set_params(self._h, p.ctypes.data, p.size), "chai_hip_mlp_set_params")

    def forward(self, X):
        """X: cuda float32 [B][inputs] -> cuda float32 [B]"""
        import torch
        out = torch.empty(X.shape[0], dtype=torch.float32, device=X.device)
        check(_bind().chai_hip_mlp_forward(self._h, X.data_ptr(), X.shape[0], out.data_ptr(), _stream()),
              "chai_hip_mlp_forward")
        return out

    def backward(self, X, Y):
        import torch
        g = torch.empty(self.num_params, dtype=torch.float32, device=X.device)
        check(_bind().chai_hip_mlp_backward(self._h, X.data_ptr(), Y.data_ptr(), X.shape[0], g.data_ptr(),
                                            _stream()), "chai_hip_mlp_backward")
        return g

    def update(self, grad, lr, adam=True):
        check(_bind().chai_hip_mlp_update(self._h, grad.data_ptr(), int(adam), lr, _stream()), "chai_hip_mlp_update")

    def train_step(self, X, Y, lr, adam=True):
        check(_bind().chai_hip_mlp_train_step(self._h, X.data_ptr(), Y.data_ptr(), X.shape[0], int(adam), lr,
                                              _stream()), "chai_hip_mlp_train_step")

    def fit(self, X, Y, epochs=300, batch=256, lr=2e-3, seed=0):
        """Adam on minibatches of the rows of (X, Y) (cuda tensors)."""
        import torch
        g = torch.Generator(device="cpu").manual_seed(seed)
        n = X.shape[0]
        for ep in range(epochs):
            perm = torch.randperm(n, generator=g).to(X.device)
            Xs, Ys = X[perm].contiguous(), Y[perm].contiguous()
            for i in range(0, n - batch + 1, batch):
                self.train_step(Xs[i:i + batch], Ys[i:i + batch], lr)
        torch.cuda.synchronize()

    def release(self):
        if getattr(self, "_h", None):
            _bind().chai_hip_mlp_release(self._h)
            self._h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


# ------------------------------------------------------------------- shapes
def random_shapes(n, seed, lo=64, hi=16384, step=16):
    """M, N, K log-uniform in [lo, hi] rounded to `step`; random transposes and
    an f32 or i32 A operand (the mixed int x float GEMM)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        mnk = np.exp(rng.uniform(np.log(lo), np.log(hi), 3))
        M, N, K = (int(max(lo, min(hi, round(v / step) * step))) for v in mnk)
        out.append((M, N, K, int(rng.integers(2)), int(rng.integers(2)), [F32, I32][int(rng.random() < 0.25)]))
    return out


def desc_of(shape, variant=-1, family=GEMM_BF16X3):
    """The GEMM descriptor of a random shape in a family: f32 (or i32 A) x f32
    for the BF16X3, BF16X6 and FMA tiles, f64 x f64 for the FP64 tiles, and for
    GEMV op(A) (M x K) times a K-vector."""
    M, N, K, ta, tb, pa = shape
    d = GemmDesc()
    d.prec_a, d.prec_b, d.prec_c = pa, F32, F32
    if family == GEMM_F64:
        d.prec_a, d.prec_b, d.prec_c = F64, F64, F64
    if family == GEMM_GEMV:
        N, tb = 1, 0
    d.trans_a, d.trans_b = ta, tb
    d.M, d.N, d.K = M, N, K
    d.lda = M if ta else K
    d.ldb = K if tb else N
    d.ldc = N
    d.batch = 1
    d.mode = family
    d.variant = variant
    return d


def num_variants(family):
    return len(gemm_variants(family))


def features(shape, variant, family=GEMM_BF16X3):
    f = np.empty(NFEAT, dtype=np.float32)
    d = desc_of(shape, family=family)
    check(_bind().chai_hip_gemm_features(ctypes.byref(d), variant, f.ctypes.data), "chai_hip_gemm_features")
    return f


class Workspace:
    """Operand buffers big enough for every shape up to hi x hi (uniform data),
    in f32, i32 and f64."""

    def __init__(self, hi=16384, f64=True):
        import torch
        from . import rng_fill
        init()
        n = hi * hi
        self.A = torch.empty(n, dtype=torch.float32, device="cuda")
        self.Ai = torch.empty(n, dtype=torch.int32, device="cuda")
        self.B = torch.empty(n, dtype=torch.float32, device="cuda")
        self.C = torch.empty(n, dtype=torch.float32, device="cuda")
        rng_fill(self.A, seed=42, limits=(-1.0, 1.0))
        rng_fill(self.B, seed=43, limits=(-1.0, 1.0))
        self.Ai.copy_((self.A * 127).to(torch.int32))
        self.Ad = self.Bd = self.Cd = None
        if f64:
            self.Ad = self.A.to(torch.float64)
            self.Bd = self.B.to(torch.float64)
            self.Cd = torch.empty(n, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()

    def ptrs(self, shape, family=GEMM_BF16X3):
        if family == GEMM_F64:
            return self.Ad.data_ptr(), self.Bd.data_ptr(), self.Cd.data_ptr()
        pa = shape[5]
        return (self.Ai if pa == I32 else self.A).data_ptr(), self.B.data_ptr(), self.C.data_ptr()


def time_variants(ws, shape, trials=3, family=GEMM_BF16X3):
    """core ms of every variant of the family (the exhaustive search) and its
    wall time (s)."""
    import torch
    L = _bind()
    nv = num_variants(family)
    a, b, c = ws.ptrs(shape, family)
    ms = np.empty(nv)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for v in range(nv):
        d = desc_of(shape, v, family)
        out = ctypes.c_float()
        check(L.chai_hip_gemm_time_variant(ctypes.byref(d), a, b, c, v, trials, ctypes.byref(out), _stream()),
              "chai_hip_gemm_time_variant")
        ms[v] = out.value
    torch.cuda.synchronize()
    return ms, time.perf_counter() - t0


def predict(model, shape, family=GEMM_BF16X3):
    """(variant, per-variant model outputs, wall seconds) of one inference."""
    L = _bind()
    nv = num_variants(family)
    d = desc_of(shape, family=family)
    v = ctypes.c_int()
    pred = np.empty(nv, dtype=np.float32)
    t0 = time.perf_counter()
    check(L.chai_hip_gemm_predict_variant(ctypes.byref(d), model._h, ctypes.byref(v), pred.ctypes.data,
                                          _stream()), "chai_hip_gemm_predict_variant")
    return v.value, pred, time.perf_counter() - t0


def dataset(shapes, times, family=GEMM_BF16X3):
    """rows (shape, variant) -> features, target log(core ms)"""
    X, Y = [], []
    for s, ms in zip(shapes, times):
        for v, t in enumerate(ms):
            if t > 0:
                X.append(features(s, v, family))
                Y.append(np.log(t))
    return np.array(X, dtype=np.float32).reshape(-1, NFEAT), np.array(Y, dtype=np.float32)


def regret(choice, times):
    """1 - perf(choice) / perf(best) per shape (perf = 1 / time)"""
    t = np.asarray(times)
    best = t.min(axis=1)
    picked = t[np.arange(len(t)), np.asarray(choice)]
    return 1.0 - best / picked


def save_model(model, path, y_mean=0.0, y_std=1.0):
    """Parameters (+ the target normalisation) as a numpy .npz (no pickle)."""
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.savez(path, params=model.params(), inputs=np.int32(model.inputs), y_mean=np.float64(y_mean),
             y_std=np.float64(y_std))


def load_model(path):
    with np.load(path, allow_pickle=False) as z:
        m = MLP(int(z["inputs"]), seed=0)
        m.set_params(z["params"])
        m.y_mean, m.y_std = float(z["y_mean"]), float(z["y_std"])
    return m
