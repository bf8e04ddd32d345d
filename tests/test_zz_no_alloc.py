"""GPU: chai_hip_reduce_noalloc (partials in a fixed device array, no
allocation) against chai_hip_reduce (partials allocated per call), called
through the C ABI.  Both run the same kernels with the same block counts, so
results must be identical bit for bit; two calls in a row, the second on a
smaller input, show that the shared array carries nothing over."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SUM, MEAN = 132, 131
U32, I32, F32, F64 = 0, 1, 4, 8


def _setup(L):
    vp, sz, i = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
    L.chai_hip_reduce_noalloc.argtypes = [i, i, vp, sz, sz, vp, vp, vp]
    L.chai_hip_reduce.argtypes = [i, i, vp, sz, sz, vp, vp, vp]


def _run(chai, fn, op, prec, x, n, slots):
    import torch
    out = torch.full((slots,), -7, dtype=x.dtype, device=x.device)
    out64 = torch.full((slots,), -7.0, dtype=torch.float64, device=x.device)
    rc = fn(op, prec, ctypes.c_void_p(x.data_ptr()), n, slots, ctypes.c_void_p(out.data_ptr()),
            ctypes.c_void_p(out64.data_ptr()), chai._stream())
    chai.check(rc, fn.__name__)
    torch.cuda.synchronize()
    return out.cpu(), out64.cpu()


def _input(prec, count, seed):
    import torch
    g = np.random.default_rng(seed)
    if prec == F32:
        a = g.uniform(-1, 1, count).astype(np.float32)
    elif prec == F64:
        a = g.uniform(-1, 1, count)
    elif prec == I32:
        a = g.integers(-2**31, 2**31, count, dtype=np.int64).astype(np.int32)
    else:
        a = g.integers(0, 2**31, count, dtype=np.int64).astype(np.int32)   # u32 words < 2^31
    return torch.from_numpy(a).cuda()


@pytest.mark.parametrize("prec", [F32, F64, I32, U32])
@pytest.mark.parametrize("op", [SUM, MEAN])
def test_noalloc_equals_allocating_reduce_twice(chai, prec, op):
    L = chai.lib()
    _setup(L)
    # first call: large input with one slot (2048 partials); second: smaller, several slots
    for n, slots, seed in [(1 << 22, 1, 1), (3001, 5, 2)]:
        print("reduce_noalloc prec=%d op=%d n=%d slots=%d" % (prec, op, n, slots), flush=True)
        x = _input(prec, n * slots, seed)
        a, a64 = _run(chai, L.chai_hip_reduce_noalloc, op, prec, x, n, slots)
        b, b64 = _run(chai, L.chai_hip_reduce, op, prec, x, n, slots)
        assert np.array_equal(a.numpy().view(np.uint8), b.numpy().view(np.uint8))
        assert np.array_equal(a64.numpy().view(np.uint64), b64.numpy().view(np.uint64))
        # and against a host sum of the same words in FP64 / int64
        h = x.cpu().numpy().reshape(slots, n)
        if prec in (I32, U32):
            ref = h.astype(np.int64).sum(axis=1).astype(np.float64)
            if op == MEAN:
                ref = ref / n
            assert np.array_equal(a64.numpy(), ref)
        else:
            ref = h.astype(np.float64).sum(axis=1)
            if op == MEAN:
                ref = ref / n
            assert np.allclose(a64.numpy(), ref, rtol=0, atol=1e-9 * n)


def test_noalloc_many_slots_and_limit(chai):
    import torch
    L = chai.lib()
    _setup(L)
    # 256 slots of 2^20 elements: 256 blocks per slot fill the 65536 partials exactly
    n = 1 << 20
    x = _input(F32, 256 * n, 3)
    a, a64 = _run(chai, L.chai_hip_reduce_noalloc, SUM, F32, x, n, 256)
    b, b64 = _run(chai, L.chai_hip_reduce, SUM, F32, x, n, 256)
    assert np.array_equal(a64.numpy().view(np.uint64), b64.numpy().view(np.uint64))
    del x
    # 300 such slots need 76800 partials: rejected on the host before any launch
    out = torch.empty(300, dtype=torch.float32, device="cuda")
    y = torch.zeros(300 * n, dtype=torch.float32, device="cuda")
    rc = L.chai_hip_reduce_noalloc(SUM, F32, ctypes.c_void_p(y.data_ptr()), n, 300,
                                   ctypes.c_void_p(out.data_ptr()), None, chai._stream())
    assert rc == -1
    assert b"scratch" in L.chai_hip_last_error()
    # the same slots with short rows fit (18 blocks each)
    z = _input(F32, 300 * 70000, 4)
    a, a64 = _run(chai, L.chai_hip_reduce_noalloc, SUM, F32, z, 70000, 300)
    b, b64 = _run(chai, L.chai_hip_reduce, SUM, F32, z, 70000, 300)
    assert np.array_equal(a64.numpy().view(np.uint64), b64.numpy().view(np.uint64))
