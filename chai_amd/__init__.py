"""chai_amd -- Python binding of the MI355X back end for Chai (libchai_mi355x.so).

The library is a C ABI (include/chai_mi355x.h); this module loads the in-tree
build with ctypes and offers thin typed wrappers that accept torch tensors
(device memory and streams come from PyTorch-ROCm; the compute is the
library's own HIP kernels).  There is no CPU fallback: if the library is not
built, or no gfx950 device is present, every call raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libchai_mi355x.so")

# PrecType codes (misc/PrecType.hpp:18-21)
U32, I32, F32, F64 = 0, 1, 4, 8
PHILOX, THREEFRY = 1230, 1231
OP_DOT, OP_MEAN, OP_SUM = 130, 131, 132
GEMM_AUTO, GEMM_FMA, GEMM_BF16X3, GEMM_F64, GEMM_INT, GEMM_BF16X6, GEMM_GEMV = 0, 1, 2, 3, 4, 5, 6

_lib = None
_inited = False


class ChaiError(RuntimeError):
    pass


class GemmDesc(ctypes.Structure):
    _fields_ = [
        ("prec_a", ctypes.c_int), ("prec_b", ctypes.c_int), ("prec_c", ctypes.c_int),
        ("trans_a", ctypes.c_int), ("trans_b", ctypes.c_int),
        ("M", ctypes.c_size_t), ("N", ctypes.c_size_t), ("K", ctypes.c_size_t),
        ("lda", ctypes.c_size_t), ("ldb", ctypes.c_size_t), ("ldc", ctypes.c_size_t),
        ("general", ctypes.c_int), ("alpha", ctypes.c_double), ("beta", ctypes.c_double),
        ("batch", ctypes.c_size_t), ("same_a", ctypes.c_int), ("same_b", ctypes.c_int),
        ("stride_a", ctypes.c_size_t), ("stride_b", ctypes.c_size_t), ("stride_c", ctypes.c_size_t),
        ("mode", ctypes.c_int), ("variant", ctypes.c_int),
    ]


def lib():
    """Load libchai_mi355x.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ChaiError("libchai_mi355x.so is not built; run `python -m chai_amd.build` "
                            "(or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        vp, sz, i, d = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_double
        L.chai_hip_last_error.restype = ctypes.c_char_p
        L.chai_hip_init.argtypes = [i]
        L.chai_hip_device_info.argtypes = [ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i),
                                           ctypes.c_char_p, sz]
        L.chai_hip_alloc.argtypes = [sz, ctypes.POINTER(vp)]
        L.chai_hip_free.argtypes = [vp]
        L.chai_hip_h2d.argtypes = [vp, vp, sz, vp]
        L.chai_hip_d2h.argtypes = [vp, vp, sz, vp]
        L.chai_hip_memset.argtypes = [vp, i, sz, vp]
        L.chai_hip_sync.argtypes = [vp]
        L.chai_hip_timer_create.argtypes = [ctypes.POINTER(vp)]
        L.chai_hip_timer_start.argtypes = [vp, vp]
        L.chai_hip_timer_stop.argtypes = [vp, vp]
        L.chai_hip_timer_elapsed_ms.argtypes = [vp, ctypes.POINTER(ctypes.c_float)]
        L.chai_hip_timer_destroy.argtypes = [vp]
        L.chai_hip_rng_fill.argtypes = [i, i, i, ctypes.c_uint64, sz, sz, sz, i, i, d, d, vp, vp]
        L.chai_hip_rng_check_normals.argtypes = [ctypes.POINTER(ctypes.c_uint64), vp]
        L.chai_hip_reduce.argtypes = [i, i, vp, sz, sz, vp, vp, vp]
        L.chai_hip_gemm.argtypes = [ctypes.POINTER(GemmDesc), vp, vp, vp, vp]
        L.chai_hip_gemm_num_variants.argtypes = [i, ctypes.POINTER(i)]
        L.chai_hip_gemm_variant_params.argtypes = [i, i, ctypes.POINTER(i)]
        L.chai_hip_profile_enable.argtypes = [i]
        L.chai_hip_gemm_last_timing.argtypes = [ctypes.POINTER(ctypes.c_float),
                                                ctypes.POINTER(ctypes.c_float)]
        L.chai_hip_gemm_profile_mean.argtypes = [i, ctypes.POINTER(ctypes.c_float),
                                                 ctypes.POINTER(ctypes.c_float)]
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != 0:
        msg = lib().chai_hip_last_error().decode(errors="replace")
        raise ChaiError("%s failed (%d): %s" % (what, rc, msg))


def init(device=0):
    """chai_hip_init on `device` (idempotent)."""
    global _inited
    if not _inited:
        check(lib().chai_hip_init(device), "chai_hip_init")
        _inited = True


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def prec_of(dtype):
    import torch
    return {torch.float32: F32, torch.float64: F64, torch.int32: I32,
            getattr(torch, "uint32", None): U32}[dtype]


def dtype_of(prec):
    import torch
    return {F32: torch.float32, F64: torch.float64, I32: torch.int32,
            U32: getattr(torch, "uint32", torch.int32)}[prec]


# ----------------------------------------------------------------------- RNG
def rng_fill(out, variant=PHILOX, seed=0, vec_len=0, normal=False, limits=None,
             W=None, H=None, slots=1, prec=None):
    """Fill `out` (a cuda tensor viewed as [slots][H][W]) with Chai's RNG stream."""
    init()
    if W is None:
        W = out.shape[-1]
    if H is None:
        H = out.numel() // (W * slots)
    p = prec if prec is not None else prec_of(out.dtype)
    lo, hi = limits if limits is not None else (0.0, 0.0)
    check(lib().chai_hip_rng_fill(variant, p, vec_len, seed, W, H, slots, int(normal),
                                  int(limits is not None), lo, hi, _ptr(out), _stream()),
          "chai_hip_rng_fill")
    return out


def rng_check_normals():
    """Run every 32-bit stream word through the f32 normals' short log / sqrt
    forms and the IEEE ones; returns (log mismatches, sqrt mismatches)."""
    init()
    out = (ctypes.c_uint64 * 2)()
    check(lib().chai_hip_rng_check_normals(out, _stream()), "chai_hip_rng_check_normals")
    return int(out[0]), int(out[1])


# --------------------------------------------------------------- reductions
def reduce(x, op="sum", slots=1, prec=None):
    """sum / mean of each of `slots` equal parts of x; returns (out, out_f64)."""
    import torch
    init()
    n = x.numel() // slots
    out = torch.empty(slots, dtype=x.dtype, device=x.device)
    out64 = torch.empty(slots, dtype=torch.float64, device=x.device)
    code = OP_MEAN if op == "mean" else OP_SUM
    p = prec if prec is not None else prec_of(x.dtype)
    check(lib().chai_hip_reduce(code, p, _ptr(x), n, slots, _ptr(out), _ptr(out64), _stream()),
          "chai_hip_reduce")
    return out, out64


# --------------------------------------------------------------------- GEMM
def gemm_variants(mode=GEMM_BF16X3):
    """[(BM, BN, BK, waves_m, waves_n), ...] of a kernel family's tile variants."""
    n = ctypes.c_int()
    check(lib().chai_hip_gemm_num_variants(mode, ctypes.byref(n)), "gemm_num_variants")
    out = []
    for v in range(n.value):
        p = (ctypes.c_int * 5)()
        if lib().chai_hip_gemm_variant_params(mode, v, p) == 0:
            out.append(tuple(p))
    return out


class Epilogue(ctypes.Structure):
    """chai_epilogue: kind 1 = max(c * scale + shift, lo) in f32."""
    _fields_ = [("kind", ctypes.c_int), ("scale", ctypes.c_float), ("shift", ctypes.c_float),
                ("lo", ctypes.c_float)]


def epilogue(scale, shift, lo):
    return Epilogue(1, scale, shift, lo)


def gemm(A, B, C=None, trans_a=False, trans_b=False, mode=GEMM_AUTO, alpha=1.0, beta=0.0,
         general=False, out_dtype=None, prec_a=None, prec_b=None, variant=-1, epi=None, prec_c=None):
    """C = op(A) op(B) (or alpha op(A) op(B) + beta C) on 2-D / 3-D (batched) tensors;
    `epi` (an Epilogue) is fused into the store."""
    import torch
    init()
    batched = A.dim() == 3 or B.dim() == 3
    nb = max(A.shape[0] if A.dim() == 3 else 1, B.shape[0] if B.dim() == 3 else 1)
    a2 = A.shape[-2:]
    b2 = B.shape[-2:]
    M, K = (a2[1], a2[0]) if trans_a else (a2[0], a2[1])
    K2, N = (b2[1], b2[0]) if trans_b else (b2[0], b2[1])
    if K != K2:
        raise ValueError("inner dimensions differ: %d vs %d" % (K, K2))
    pa = prec_a if prec_a is not None else prec_of(A.dtype)
    pb = prec_b if prec_b is not None else prec_of(B.dtype)
    if C is None:
        if out_dtype is None:
            out_dtype = torch.float64 if F64 in (pa, pb) else torch.float32
        C = torch.empty((nb, M, N) if batched else (M, N), dtype=out_dtype, device=A.device)
    d = GemmDesc()
    d.prec_a, d.prec_b, d.prec_c = pa, pb, prec_c if prec_c is not None else prec_of(C.dtype)
    d.trans_a, d.trans_b = int(trans_a), int(trans_b)
    d.M, d.N, d.K = M, N, K
    d.lda, d.ldb, d.ldc = A.stride(-2), B.stride(-2), C.stride(-2)
    d.general = int(general)
    d.alpha, d.beta = alpha, beta
    d.batch = nb if batched else 1
    d.same_a = int(A.dim() == 2 and batched)
    d.same_b = int(B.dim() == 2 and batched)
    d.stride_a = A.stride(0) if A.dim() == 3 else 0
    d.stride_b = B.stride(0) if B.dim() == 3 else 0
    d.stride_c = C.stride(0) if C.dim() == 3 else 0
    d.mode = mode
    d.variant = variant
    if epi is None:
        check(lib().chai_hip_gemm(ctypes.byref(d), _ptr(A), _ptr(B), _ptr(C), _stream()), "chai_hip_gemm")
    else:
        L = lib()
        L.chai_hip_gemm_epilogue.argtypes = [ctypes.POINTER(GemmDesc)] + [ctypes.c_void_p] * 3 + [
            ctypes.POINTER(Epilogue), ctypes.c_void_p]
        check(L.chai_hip_gemm_epilogue(ctypes.byref(d), _ptr(A), _ptr(B), _ptr(C), ctypes.byref(epi),
                                       _stream()), "chai_hip_gemm_epilogue")
    return C


def gemv(A, x, y=None, trans_a=False, general=False, alpha=1.0, beta=0.0, prec_a=None, prec_x=None):
    """y = op(A) x (or alpha op(A) x + beta y); A is M x N row-major (2-D tensor)."""
    import torch
    init()
    M, N = A.shape
    pa = prec_a if prec_a is not None else prec_of(A.dtype)
    px = prec_x if prec_x is not None else prec_of(x.dtype)
    if pa <= I32 and px <= I32:
        py = max(pa, px)                 # integer products: the promoted integer type
    else:
        py = F64 if F64 in (pa, px) else F32
    if y is None:
        y = torch.zeros(N if trans_a else M, dtype=dtype_of(py), device=A.device)
    L = lib()
    L.chai_hip_gemv.argtypes = [ctypes.c_int] * 4 + [ctypes.c_size_t] * 2 + [
        ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
        ctypes.c_double, ctypes.c_double, ctypes.c_size_t, ctypes.c_int, ctypes.c_size_t,
        ctypes.c_size_t, ctypes.c_size_t, ctypes.c_void_p]
    check(L.chai_hip_gemv(pa, px, prec_of(y.dtype), int(trans_a), M, N, _ptr(A), A.stride(0), _ptr(x),
                          _ptr(y), int(general), alpha, beta, 1, 0, 0, 0, 0, _stream()), "chai_hip_gemv")
    return y


# ------------------------------------------------------------------ tuning
def memo_open(path, device=0):
    """Load + journal to `path` (autotune/AutoTuneMemo semantics)."""
    L = lib()
    L.chai_hip_memo_open.argtypes = [ctypes.c_char_p, ctypes.c_int]
    check(L.chai_hip_memo_open(str(path).encode(), device), "memo_open")


def memo_close():
    check(lib().chai_hip_memo_close(), "memo_close")


def memo_save(path):
    L = lib()
    L.chai_hip_memo_save.argtypes = [ctypes.c_char_p]
    check(L.chai_hip_memo_save(str(path).encode()), "memo_save")


def memo_record(kernel, value):
    L = lib()
    L.chai_hip_memo_record.argtypes = [ctypes.c_char_p, ctypes.c_long]
    check(L.chai_hip_memo_record(kernel.encode(), value), "memo_record")


def memo_status(kernel):
    """(status, us): 1 good / 0 unknown / -1 blacklisted."""
    L = lib()
    L.chai_hip_memo_kernel_status.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int),
                                              ctypes.POINTER(ctypes.c_long)]
    s, us = ctypes.c_int(), ctypes.c_long()
    check(L.chai_hip_memo_kernel_status(kernel.encode(), ctypes.byref(s), ctypes.byref(us)), "memo_status")
    return s.value, us.value


def _desc(A, B, C, trans_a=False, trans_b=False, mode=GEMM_AUTO, variant=-1):
    M, K = (A.shape[1], A.shape[0]) if trans_a else (A.shape[0], A.shape[1])
    N = B.shape[0] if trans_b else B.shape[1]
    d = GemmDesc()
    d.prec_a, d.prec_b, d.prec_c = prec_of(A.dtype), prec_of(B.dtype), prec_of(C.dtype)
    d.trans_a, d.trans_b = int(trans_a), int(trans_b)
    d.M, d.N, d.K = M, N, K
    d.lda, d.ldb, d.ldc = A.stride(0), B.stride(0), C.stride(0)
    d.batch = 1
    d.mode = mode
    d.variant = variant
    return d


def gemm_time_variant(A, B, C, variant, mode=GEMM_BF16X3, trials=3):
    """core-kernel ms of one tile variant (journaled); -1 if blacklisted."""
    init()
    d = _desc(A, B, C, mode=mode)
    L = lib()
    L.chai_hip_gemm_time_variant.argtypes = [ctypes.POINTER(GemmDesc), ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_float), ctypes.c_void_p]
    ms = ctypes.c_float()
    check(L.chai_hip_gemm_time_variant(ctypes.byref(d), _ptr(A), _ptr(B), _ptr(C), variant, trials,
                                       ctypes.byref(ms), _stream()), "gemm_time_variant")
    return ms.value


def gemm_tune(A, B, C, mode=GEMM_BF16X3, trials=3):
    """exhaustive variant search for this shape; returns (variant, core ms)."""
    init()
    d = _desc(A, B, C, mode=mode)
    L = lib()
    L.chai_hip_gemm_tune.argtypes = [ctypes.POINTER(GemmDesc), ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                     ctypes.POINTER(ctypes.c_float), ctypes.c_void_p]
    v, ms = ctypes.c_int(), ctypes.c_float()
    check(L.chai_hip_gemm_tune(ctypes.byref(d), _ptr(A), _ptr(B), _ptr(C), trials, ctypes.byref(v),
                               ctypes.byref(ms), _stream()), "gemm_tune")
    return v.value, ms.value


def gemm_set_variant(A, B, C, variant, mode=GEMM_BF16X3):
    """record `variant` for this shape (used when gemm() gets variant=-1)."""
    d = _desc(A, B, C, mode=mode)
    L = lib()
    L.chai_hip_gemm_set_variant.argtypes = [ctypes.POINTER(GemmDesc), ctypes.c_int]
    check(L.chai_hip_gemm_set_variant(ctypes.byref(d), variant), "gemm_set_variant")


def profile(on=True):
    """Enable per-kernel hipEvent timing inside chai_hip_gemm."""
    check(lib().chai_hip_profile_enable(int(on)), "profile_enable")


def gemm_last_timing():
    """(pre-pass ms, core kernel ms) of this thread's last profiled GEMM."""
    a, b = ctypes.c_float(), ctypes.c_float()
    check(lib().chai_hip_gemm_last_timing(ctypes.byref(a), ctypes.byref(b)), "gemm_last_timing")
    return a.value, b.value


def gemm_profile_mean(count):
    """(mean pre-pass ms, mean core kernel ms) over this thread's last `count`
    profiled GEMMs (at most 64 kept since profiling was enabled)."""
    a, b = ctypes.c_float(), ctypes.c_float()
    check(lib().chai_hip_gemm_profile_mean(int(count), ctypes.byref(a), ctypes.byref(b)), "gemm_profile_mean")
    return a.value, b.value


class Timer:
    """hipEvent timer on the current torch stream (the stream the library launches on)."""

    def __init__(self):
        init()
        self.h = ctypes.c_void_p()
        check(lib().chai_hip_timer_create(ctypes.byref(self.h)), "timer_create")

    def start(self):
        check(lib().chai_hip_timer_start(self.h, _stream()), "timer_start")

    def stop(self):
        check(lib().chai_hip_timer_stop(self.h, _stream()), "timer_stop")

    def ms(self):
        v = ctypes.c_float()
        check(lib().chai_hip_timer_elapsed_ms(self.h, ctypes.byref(v)), "timer_elapsed")
        return v.value

    def __del__(self):
        try:
            lib().chai_hip_timer_destroy(self.h)
        except Exception:
            pass
