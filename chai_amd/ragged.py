"""Ragged GEMM batches and the thread-safe batcher (chai_hip_ragged_* /
chai_hip_batcher_*).

The reference gathers traces with an identical hash from many host threads
into one VectorTrace (runtime/Scheduler.cpp:48-307, common/VectorTrace.cpp:
9-100); instances of different sizes are never batched there.  Here any mix
of f32 GEMMs C = A B (+ the fused epilogue max(C*scale+shift, lo)) runs as one
split pre-pass and one persistent MFMA launch, optionally captured in a
hipGraph.
"""
import ctypes

from . import ChaiError, Epilogue, _stream, check, init, lib

_vp, _sz = ctypes.c_void_p, ctypes.c_size_t


class RaggedItem(ctypes.Structure):
    _fields_ = [("A", _vp), ("B", _vp), ("C", _vp),
                ("M", _sz), ("N", _sz), ("K", _sz),
                ("lda", _sz), ("ldb", _sz), ("ldc", _sz)]


def item(A, B, C):
    """RaggedItem of 2-D f32 cuda tensors (row-major; row strides from the tensors)."""
    import torch
    for t in (A, B, C):
        if t.dtype != torch.float32 or t.dim() != 2 or t.stride(1) != 1:
            raise ValueError("ragged items are 2-D row-major float32 tensors")
    M, K = A.shape
    K2, N = B.shape
    if K != K2 or tuple(C.shape) != (M, N):
        raise ValueError("shapes do not chain: %s x %s -> %s" % (tuple(A.shape), tuple(B.shape), tuple(C.shape)))
    return RaggedItem(A.data_ptr(), B.data_ptr(), C.data_ptr(), M, N, K, A.stride(0), B.stride(0), C.stride(0))


def _bind():
    L = lib()
    if getattr(L, "_ragged_bound", False):
        return L
    P = ctypes.POINTER
    L.chai_hip_ragged_plan.argtypes = [P(RaggedItem), _sz, P(Epilogue), ctypes.c_int, P(_vp)]
    L.chai_hip_ragged_run.argtypes = [_vp, ctypes.c_int, _vp]
    L.chai_hip_ragged_info.argtypes = [_vp, P(ctypes.c_uint64)]
    L.chai_hip_ragged_last_timing.argtypes = [_vp, P(ctypes.c_float), P(ctypes.c_float)]
    L.chai_hip_ragged_release.argtypes = [_vp]
    L.chai_hip_ragged_tune.argtypes = [P(RaggedItem), _sz, P(Epilogue), ctypes.c_int, P(ctypes.c_int),
                                       P(ctypes.c_float), _vp]
    L.chai_hip_batcher_create.argtypes = [_sz, P(Epilogue), ctypes.c_int, ctypes.c_int, ctypes.c_int, P(_vp)]
    L.chai_hip_batcher_submit.argtypes = [_vp, P(RaggedItem), _sz, _vp]
    L.chai_hip_batcher_flush.argtypes = [_vp]
    L.chai_hip_batcher_stats.argtypes = [_vp, P(ctypes.c_uint64)]
    L.chai_hip_batcher_destroy.argtypes = [_vp]
    L._ragged_bound = True
    return L


def _items_array(items):
    arr = (RaggedItem * max(len(items), 1))(*items)
    return arr, len(items)


class Plan:
    """A compiled ragged batch: owns its workspace (and graph once replayed)."""

    def __init__(self, items, epi=None, variant=-1):
        init()
        L = _bind()
        self._items = list(items)
        arr, n = _items_array(self._items)
        h = _vp()
        check(L.chai_hip_ragged_plan(arr, n, ctypes.byref(epi) if epi is not None else None, variant,
                                     ctypes.byref(h)), "chai_hip_ragged_plan")
        self._h = h

    def run(self, graph=False):
        check(_bind().chai_hip_ragged_run(self._h, int(graph), _stream()), "chai_hip_ragged_run")

    def info(self):
        out = (ctypes.c_uint64 * 7)()
        check(_bind().chai_hip_ragged_info(self._h, out), "chai_hip_ragged_info")
        keys = ("tiles", "blocks", "workspace_bytes", "useful_flops", "padded_flops", "variant", "split_blocks")
        return dict(zip(keys, list(out)))

    def last_timing(self):
        a, b = ctypes.c_float(), ctypes.c_float()
        check(_bind().chai_hip_ragged_last_timing(self._h, ctypes.byref(a), ctypes.byref(b)),
              "chai_hip_ragged_last_timing")
        return a.value, b.value

    def release(self):
        if getattr(self, "_h", None):
            _bind().chai_hip_ragged_release(self._h)
            self._h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def num_variants():
    """Number of ragged tile variants (0..7 panel tiles after a split pre-pass, 8..9 f32-fed)."""
    n = ctypes.c_int()
    check(lib().chai_hip_ragged_num_variants(ctypes.byref(n)), "chai_hip_ragged_num_variants")
    return n.value


def tune(items, epi=None, trials=3):
    """Time every tile variant on this batch; returns (best_variant, best_ms)."""
    init()
    L = _bind()
    arr, n = _items_array(list(items))
    v, ms = ctypes.c_int(), ctypes.c_float()
    check(L.chai_hip_ragged_tune(arr, n, ctypes.byref(epi) if epi is not None else None, trials,
                                 ctypes.byref(v), ctypes.byref(ms), _stream()), "chai_hip_ragged_tune")
    return v.value, ms.value


class Batcher:
    """Thread-safe gathering of ragged items; submit() blocks until the batch
    holding the item has run (ctypes releases the GIL during the call)."""

    def __init__(self, expected, epi=None, variant=-1, graph=True, timeout_ms=0):
        init()
        L = _bind()
        h = _vp()
        check(L.chai_hip_batcher_create(expected, ctypes.byref(epi) if epi is not None else None, variant,
                                        int(graph), timeout_ms, ctypes.byref(h)), "chai_hip_batcher_create")
        self._h = h

    def submit(self, items, stream=None):
        """Queue one item or a list of items (one thread's trace) and wait for
        them.  The batch runs after the work queued so far on `stream` (a
        torch stream; default: the calling thread's current stream), which
        is expected to produce the items' operands."""
        if isinstance(items, RaggedItem):
            items = [items]
        arr, n = _items_array(list(items))
        s = _vp(stream.cuda_stream) if stream is not None else _stream()
        check(_bind().chai_hip_batcher_submit(self._h, arr, n, s), "chai_hip_batcher_submit")

    def flush(self):
        check(_bind().chai_hip_batcher_flush(self._h), "chai_hip_batcher_flush")

    def stats(self):
        out = (ctypes.c_uint64 * 5)()
        check(_bind().chai_hip_batcher_stats(self._h, out), "chai_hip_batcher_stats")
        s = list(out)
        return {"batches": s[0], "items": s[1], "plans": s[2], "replays": s[3],
                "variant": ctypes.c_int64(s[4]).value}

    def close(self):
        if getattr(self, "_h", None):
            check(_bind().chai_hip_batcher_destroy(self._h), "chai_hip_batcher_destroy")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except (ChaiError, Exception):
            pass
