"""Pins oracle pieces against the reference's own sources compiled into
oracle/_ref (misc/TEA.cpp, misc/PrecType.cpp, misc/SimpleFun*.cpp), built by
`make -C oracle ref` in the container that holds the reference tree."""
import ctypes

import numpy as np
import pytest


@pytest.fixture(scope="module")
def ref(oracle):
    from oracle import oracle as o
    L = o.ref_lib()
    if L is None:
        pytest.skip("oracle/_ref not built (the reference tree is not present here)")
    return L


def test_tea_hash_matches_reference(oracle, ref):
    rng = np.random.default_rng(0)
    for n in [0, 1, 2, 5, 17, 100]:
        w = rng.integers(0, 2 ** 63, size=n, dtype=np.uint64)
        buf = np.ascontiguousarray(w)
        assert oracle.tea_hash(buf) == ref.chai_ref_tea_hash(buf.ctypes.data_as(ctypes.c_void_p), n)


def test_pad_width_matches_reference(ref):
    for prec, vl in [(0, 4), (1, 4), (4, 4), (8, 2)]:
        for W in range(1, 20):
            expect = W if W % vl == 0 else vl * (1 + W // vl)
            assert ref.chai_ref_pad_width(prec, W) == expect
