"""Writes tests/golden/interp_fixtures.npz: inputs and outputs of the
reference's own interpreter on the reference-API programs of tests/frontend/
(fe_cases.cpp), run with the interpreter only.

The reference's library is compiled here from its sources (oracle/frontend.mk,
tests/frontend/Makefile); it never leaves this machine, so its results are
kept as data: config 1's f32 256^3 product on inputs from Chai's Philox stream
(seed 42), the integer and mixed products, and the transposed, general and
matrix-vector products.  The fixture pins the oracle's
restatement of the interpreter (tests/test_golden_interp.py) wherever the
reference tree is absent.

Run (this container, after `make -C tests/frontend`):
    python tests/golden/make_interp_fixtures.py
"""
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from tests import fe_programs as fe   # noqa: E402

CASES = ["mm_f32", "mm_shapes", "mm_int"]


def main():
    if not fe.built():
        sys.exit("reference-API programs are not built: make -C tests/frontend")
    arrays = {}
    with tempfile.TemporaryDirectory() as d:
        for case in CASES:
            outs, _ = fe.run_case(case, os.path.join(d, case))
            for name, a in fe.case_inputs(case).items():
                if case != "mm_f32":          # config 1's inputs come from the Philox stream (seed 42)
                    arrays["%s.in.%s" % (case, name)] = a
            for name, a in outs.items():
                arrays["%s.out.%s" % (case, name)] = a
    path = os.path.join(ROOT, "tests", "golden", "interp_fixtures.npz")
    np.savez_compressed(path, **arrays)
    print("wrote %s: %d arrays, %d bytes" % (path, len(arrays), os.path.getsize(path)))


if __name__ == "__main__":
    main()
