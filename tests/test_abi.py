"""CPU checks of the C ABI: the library loads here (no GPU needed to load it)
and exports every entry point that include/*.h declares."""
import ctypes
import glob
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "chai_amd", "libchai_mi355x.so")


def declared_symbols():
    syms = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(chai_hip_\w+)\s*\(", text):
            syms.add(m.group(1))
    return syms


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ["chai_hip_init", "chai_hip_rng_fill", "chai_hip_reduce", "chai_hip_gemm"]:
        assert s in syms


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail("libchai_mi355x.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = sorted(declared_symbols() - exported)
    assert not missing, missing
    lib = ctypes.CDLL(LIB)
    for s in declared_symbols():
        assert hasattr(lib, s)


def test_header_compiles_as_c():
    src = '#include "chai_mi355x.h"\nint main(void){chai_gemm_desc d; (void)d; return 0;}\n'
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only",
                        "-I", os.path.join(ROOT, "include"), "-x", "c", "-"],
                       input=src, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_python_binding_fails_loudly_without_library(tmp_path, monkeypatch):
    import chai_amd
    monkeypatch.setattr(chai_amd, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(chai_amd, "_lib", None)
    with pytest.raises(chai_amd.ChaiError):
        chai_amd.lib()
