"""The oracle's restatement of the interpreter's products against the
reference interpreter's own results (tests/golden/interp_fixtures.npz, written
by tests/golden/make_interp_fixtures.py from the reference's library compiled
in this container): bit-exact for every case -- the f32 product of config 1
(256^3, Philox inputs with seed 42, a double accumulator rounded once), the
integer products (wrapped products in the promoted type, x86 conversions of
the sum), the mixed i32 x f32 product, transposed operands, matmulG and
matrix-vector."""
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIX = np.load(os.path.join(ROOT, "tests", "golden", "interp_fixtures.npz"))


def _in(case, name):
    return FIX["%s.in.%s" % (case, name)]


def _out(case, name):
    return FIX["%s.out.%s" % (case, name)]


def _config1_inputs(oracle):
    A = oracle.rng_fill(1230, 4, 4, 42, 256, 256, limits=(-1.0, 1.0)).reshape(256, 256)
    B = oracle.rng_fill(1230, 4, 4, 42 + 65536, 256, 256, limits=(-1.0, 1.0)).reshape(256, 256)
    return A, B


def test_config1_f32_product(oracle):
    A, B = _config1_inputs(oracle)
    C = oracle.matmul_interp(A, B)
    assert np.array_equal(C.ravel().view(np.uint32), _out("mm_f32", "C").view(np.uint32))


def test_integer_and_mixed_products(oracle):
    M, K, N = 64, 48, 40
    UA, UB = _in("mm_int", "UA").reshape(M, K), _in("mm_int", "UB").reshape(K, N)
    IA, IB = _in("mm_int", "IA").reshape(M, K), _in("mm_int", "IB").reshape(K, N)
    FB = _in("mm_int", "FB").reshape(K, N)
    assert np.array_equal(oracle.matmul_interp(UA, UB).ravel(), _out("mm_int", "CUU"))
    assert np.array_equal(oracle.matmul_interp(IA, IB).ravel(), _out("mm_int", "CII"))
    assert np.array_equal(oracle.matmul_interp(UA, IB, out_dtype=np.int32).ravel(), _out("mm_int", "CUI"))
    cif = oracle.matmul_interp(IA, FB).ravel()
    assert np.array_equal(cif.view(np.uint32), _out("mm_int", "CIF").view(np.uint32))
    # the interpreter's int32 results overflow to INT_MIN, and u32 x i32 sums are nonnegative
    assert (_out("mm_int", "CII") == np.iinfo(np.int32).min).any()


def test_transposed_general_and_matrix_vector(oracle):
    M, K, N = 96, 160, 72
    A, At = _in("mm_shapes", "A").reshape(M, K), _in("mm_shapes", "At").reshape(K, M)
    B, Bt = _in("mm_shapes", "B").reshape(K, N), _in("mm_shapes", "Bt").reshape(N, K)
    C0, x = _in("mm_shapes", "C0").reshape(M, N), _in("mm_shapes", "x")
    got = {
        "C1": oracle.matmul_interp(At, B, trans_a=True),
        "C2": oracle.matmul_interp(A, Bt, trans_b=True),
        "C3": oracle.matmul_interp(At, Bt, trans_a=True, trans_b=True),
        "C4": oracle.matmul_interp(A, B, C=C0, alpha=0.75, beta=-1.5),
        "y": oracle.matvec_interp(A, x),
    }
    for name, v in got.items():
        assert np.array_equal(np.ascontiguousarray(v).ravel().view(np.uint32),
                              _out("mm_shapes", name).view(np.uint32)), name
