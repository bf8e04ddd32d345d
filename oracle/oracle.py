"""ctypes loader for the CPU oracle (oracle/_build/libchai_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "_build", "libchai_oracle.so")
_REF = os.path.join(_HERE, "_ref", "libchai_ref.so")
_lib = None

PREC_DTYPE = {0: np.uint32, 1: np.int32, 4: np.float32, 8: np.float64}
DTYPE_PREC = {np.dtype(np.uint32): 0, np.dtype(np.int32): 1,
              np.dtype(np.float32): 4, np.dtype(np.float64): 8}


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        vp, sz, i = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
        L.chai_or_rng_fill.argtypes = [i, i, i, ctypes.c_uint64, sz, sz, sz, i, i,
                                       ctypes.c_double, ctypes.c_double, vp]
        L.chai_or_log.restype = ctypes.c_double
        L.chai_or_log.argtypes = [ctypes.c_double]
        L.chai_or_logf.restype = ctypes.c_float
        L.chai_or_logf.argtypes = [ctypes.c_float]
        L.chai_or_sincos2pif.argtypes = [ctypes.c_float, ctypes.POINTER(ctypes.c_float),
                                         ctypes.POINTER(ctypes.c_float)]
        L.chai_or_sincos2pi.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_double)]
        L.chai_or_matmul_interp.argtypes = [i, i, i, i, i, sz, sz, sz, vp, sz, vp, sz, vp, sz,
                                            i, ctypes.c_double, ctypes.c_double]
        L.chai_or_gemm_f64ref.argtypes = [i, i, i, i, sz, sz, sz, vp, sz, vp, sz, vp, sz, i]
        L.chai_or_matvec_interp.argtypes = [i, i, i, i, sz, sz, vp, sz, vp, vp, i,
                                            ctypes.c_double, ctypes.c_double]
        L.chai_or_reduce.argtypes = [i, i, vp, sz, sz, vp]
        L.chai_or_dot.argtypes = [i, i, vp, vp, sz, sz, vp]
        L.chai_or_tea_hash.restype = ctypes.c_uint64
        L.chai_or_tea_hash.argtypes = [vp, sz]
        L.chai_or_trace_eval.argtypes = [vp, vp]
        L.chai_or_mlp_num_params.restype = sz
        L.chai_or_mlp_num_params.argtypes = [i]
        L.chai_or_mlp_forward.argtypes = [vp, i, vp, sz, vp]
        L.chai_or_mlp_gradient.restype = ctypes.c_double
        L.chai_or_mlp_gradient.argtypes = [vp, i, vp, vp, sz, vp]
        _lib = L
    return _lib


def ref_lib():
    """oracle/_ref/libchai_ref.so (reference misc/ sources), or None if not built."""
    if not os.path.exists(_REF):
        return None
    L = ctypes.CDLL(_REF)
    L.chai_ref_tea_hash.restype = ctypes.c_uint64
    L.chai_ref_tea_hash.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
    L.chai_ref_pad_width.restype = ctypes.c_size_t
    L.chai_ref_pad_width.argtypes = [ctypes.c_size_t, ctypes.c_size_t]
    for name in ("chai_ref_fun2_int",):
        getattr(L, name).argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
    L.chai_ref_fun1_int.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_size_t]
    L.chai_ref_fun2_f32.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_size_t]
    if hasattr(L, "chai_ref_fun3"):
        L.chai_ref_fun1_fp.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_size_t]
        L.chai_ref_fun2_fp.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_size_t]
        L.chai_ref_fun3.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 4 + [ctypes.c_size_t]
    return L


_R123 = os.path.join(_HERE, "_ref", "libchai_r123.so")


def r123_lib():
    """oracle/_ref/libchai_r123.so (the reference's emitted Random123 OpenCL
    text compiled for the host), or None if not built."""
    if not os.path.exists(_R123):
        return None
    L = ctypes.CDLL(_R123)
    vp = ctypes.c_void_p
    for name in ("chai_ref_philox4x32", "chai_ref_philox2x32", "chai_ref_threefry4x32",
                 "chai_ref_threefry2x32", "chai_ref_threefry2x64"):
        getattr(L, name).argtypes = [vp, vp, vp]
    L.chai_ref_stream_words.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, vp]
    return L


def rng_words(variant, pr, vl, vidx, seed):
    """raw words of one stream vector (pr 32: vl words; pr 64: 2 words)"""
    o = (ctypes.c_uint64 * 4)()
    L = lib()
    L.chai_or_rng_words.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                    ctypes.c_uint64, ctypes.c_void_p]
    if L.chai_or_rng_words(variant, pr, vl, vidx, seed, o) != 0:
        raise ValueError("bad stream parameters")
    return list(o)[:(2 if pr == 64 else vl)]


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def philox4x32(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr); k = (ctypes.c_uint32 * 2)(*key); o = (ctypes.c_uint32 * 4)()
    lib().chai_or_philox4x32_10(c, k, o)
    return list(o)


def philox2x32(ctr, key):
    c = (ctypes.c_uint32 * 2)(*ctr); k = (ctypes.c_uint32 * 1)(*key); o = (ctypes.c_uint32 * 2)()
    lib().chai_or_philox2x32_10(c, k, o)
    return list(o)


def philox2x64(ctr, key):
    c = (ctypes.c_uint64 * 2)(*ctr); k = (ctypes.c_uint64 * 1)(*key); o = (ctypes.c_uint64 * 2)()
    lib().chai_or_philox2x64_10(c, k, o)
    return list(o)


def threefry4x32(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr); k = (ctypes.c_uint32 * 4)(*key); o = (ctypes.c_uint32 * 4)()
    lib().chai_or_threefry4x32_20(c, k, o)
    return list(o)


def threefry2x32(ctr, key):
    c = (ctypes.c_uint32 * 2)(*ctr); k = (ctypes.c_uint32 * 2)(*key); o = (ctypes.c_uint32 * 2)()
    lib().chai_or_threefry2x32_20(c, k, o)
    return list(o)


def threefry2x64(ctr, key):
    c = (ctypes.c_uint64 * 2)(*ctr); k = (ctypes.c_uint64 * 2)(*key); o = (ctypes.c_uint64 * 2)()
    lib().chai_or_threefry2x64_20(c, k, o)
    return list(o)


def log(x):
    return lib().chai_or_log(float(x))


def logf(x):
    return lib().chai_or_logf(float(x))


def sincos2pif(u):
    s = ctypes.c_float(); c = ctypes.c_float()
    lib().chai_or_sincos2pif(float(u), ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def sincos2pi(u):
    s = ctypes.c_double(); c = ctypes.c_double()
    lib().chai_or_sincos2pi(float(u), ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def rng_fill(variant, prec, vec_len, seed, W, H=1, slots=1, normal=False,
             limits=None):
    out = np.empty((slots, H, W), dtype=PREC_DTYPE[prec])
    lo, hi = limits if limits is not None else (0.0, 0.0)
    rc = lib().chai_or_rng_fill(variant, prec, vec_len, seed, W, H, slots, int(normal),
                                int(limits is not None), lo, hi, _p(out))
    if rc != 0:
        raise ValueError("chai_or_rng_fill rejected its arguments")
    return out


def matmul_interp(A, B, trans_a=False, trans_b=False, out_dtype=None, C=None,
                  alpha=1.0, beta=0.0):
    A = np.ascontiguousarray(A); B = np.ascontiguousarray(B)
    M = A.shape[1] if trans_a else A.shape[0]
    K = A.shape[0] if trans_a else A.shape[1]
    N = B.shape[0] if trans_b else B.shape[1]
    pa, pb = DTYPE_PREC[A.dtype], DTYPE_PREC[B.dtype]
    general = C is not None
    if general:
        out = np.array(C, copy=True)
    else:
        dt = out_dtype or PREC_DTYPE[max(pa, pb)]
        out = np.zeros((M, N), dtype=dt)
    pc = DTYPE_PREC[out.dtype]
    rc = lib().chai_or_matmul_interp(pa, pb, pc, int(trans_a), int(trans_b), M, N, K,
                                     _p(A), A.shape[1], _p(B), B.shape[1], _p(out), N,
                                     int(general), alpha, beta)
    assert rc == 0
    return out


def gemm_f64ref(A, B, trans_a=False, trans_b=False, nthreads=1):
    A = np.ascontiguousarray(A); B = np.ascontiguousarray(B)
    M = A.shape[1] if trans_a else A.shape[0]
    K = A.shape[0] if trans_a else A.shape[1]
    N = B.shape[0] if trans_b else B.shape[1]
    out = np.empty((M, N), dtype=np.float64)
    rc = lib().chai_or_gemm_f64ref(DTYPE_PREC[A.dtype], DTYPE_PREC[B.dtype], int(trans_a),
                                   int(trans_b), M, N, K, _p(A), A.shape[1], _p(B), B.shape[1],
                                   _p(out), N, nthreads)
    assert rc == 0
    return out


def matvec_interp(A, x, trans_a=False, out_dtype=None, y=None, alpha=1.0, beta=0.0):
    A = np.ascontiguousarray(A); x = np.ascontiguousarray(x)
    M = A.shape[1] if trans_a else A.shape[0]
    N = A.shape[0] if trans_a else A.shape[1]
    pa, px = DTYPE_PREC[A.dtype], DTYPE_PREC[x.dtype]
    general = y is not None
    out = np.array(y, copy=True) if general else np.zeros(M, dtype=out_dtype or PREC_DTYPE[max(pa, px)])
    rc = lib().chai_or_matvec_interp(pa, px, DTYPE_PREC[out.dtype], int(trans_a), M, N, _p(A),
                                     A.shape[1], _p(x), _p(out), int(general), alpha, beta)
    assert rc == 0
    return out


def reduce(op, a):
    """op: 'sum' or 'mean'; a: [slots, n] array."""
    a = np.ascontiguousarray(a)
    a2 = a.reshape(a.shape[0], -1) if a.ndim > 1 else a.reshape(1, -1)
    out = np.empty(a2.shape[0], dtype=a.dtype)
    rc = lib().chai_or_reduce(131 if op == "mean" else 132, DTYPE_PREC[a.dtype], _p(a2),
                              a2.shape[1], a2.shape[0], _p(out))
    assert rc == 0
    return out


def trace_eval(trace, inputs):
    """Evaluate a chai_amd.trace.Trace with interpreter semantics on the host.
    inputs: {var index: numpy array}; returns {var index: numpy array} for all variables."""
    c = trace.to_c()
    bufs = []
    arrs = {}
    for i, (p, role, W, H, S) in enumerate(trace.vars):
        if i in inputs:
            a = np.ascontiguousarray(inputs[i], dtype=PREC_DTYPE[p]).reshape(S, H, W).copy()
        else:
            a = np.zeros((S, H, W), dtype=PREC_DTYPE[p])
        arrs[i] = a
        bufs.append(a.ctypes.data_as(ctypes.c_void_p))
    arr = (ctypes.c_void_p * len(bufs))(*[b.value for b in bufs])
    rc = lib().chai_or_trace_eval(ctypes.byref(c), arr)
    if rc != 0:
        raise ValueError("oracle trace evaluation failed")
    return arrs


def tea_hash(words):
    w = np.ascontiguousarray(words, dtype=np.uint64)
    return int(lib().chai_or_tea_hash(_p(w), w.size))


# ------------------------------------------------------------ tile-cost MLP
def mlp_num_params(F):
    return int(lib().chai_or_mlp_num_params(F))


def mlp_forward(params, X):
    """outputs (float64) of the 3 x 64 MLP for rows of X (f32 [B][F])"""
    p = np.ascontiguousarray(params, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32)
    out = np.empty(X.shape[0], dtype=np.float64)
    lib().chai_or_mlp_forward(_p(p), X.shape[1], _p(X), X.shape[0], _p(out))
    return out


def mlp_gradient(params, X, Y):
    """(loss, gradient) of mean (y - Y)^2 with respect to the f32 parameters, in double"""
    p = np.ascontiguousarray(params, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32)
    Y = np.ascontiguousarray(Y, dtype=np.float32)
    g = np.empty(p.size, dtype=np.float64)
    loss = lib().chai_or_mlp_gradient(_p(p), X.shape[1], _p(X), _p(Y), X.shape[0], _p(g))
    return loss, g
