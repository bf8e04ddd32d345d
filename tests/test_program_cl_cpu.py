"""CPU checks of the inline OpenCL C hook (ProgramCL, api/ProgramCL.cpp): the
text compiles for gfx950 without a GPU, kernel parameters are read from the
code object's metadata with ProgramCL's argument kinds, compile errors come
back with the compiler log, launches are checked before anything runs, and
traces with inline kernel statements compile into separate kernels."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

from chai_amd import ChaiError, F32, U32
from chai_amd import program_cl as pcl
from chai_amd import trace as tr
from tests import cl_kernels as K


def test_parameters_from_metadata():
    p = pcl.ProgramCL(K.ADD_SCALAR)
    assert p.args("add_scalar") == [(pcl.ARRAY, 8, "float*", "a"), (pcl.SCALAR, 4, "float", "s")]
    b = pcl.ProgramCL(K.BITONIC)
    assert [a[0] for a in b.args("bitonic_pass")] == [pcl.ARRAY] + [pcl.SCALAR] * 4
    assert [a[2] for a in b.args("bitonic_pass")][1:] == ["uint"] * 4
    s = pcl.ProgramCL(K.BLOCK_SUM)
    assert [a[0] for a in s.args("block_sum")] == [pcl.ARRAY, pcl.ARRAY, pcl.LOCAL, pcl.SCALAR]
    am = pcl.ProgramCL(K.AMERICAN, "-DSTEPS=100 -DGROUP=5")
    assert [a[2] for a in am.args("american_call")] == ["double*", "double*", "double*", "double", "double", "double"]


def test_list_of_strings_is_joined():
    half = len(K.ADD_SCALAR) // 2
    p = pcl.ProgramCL([K.ADD_SCALAR[:half], K.ADD_SCALAR[half:]])
    assert p.args("add_scalar")[0][2] == "float*"


def test_same_text_is_cached():
    assert pcl.ProgramCL(K.ADD_SCALAR).h.value == pcl.ProgramCL(K.ADD_SCALAR).h.value
    assert pcl.ProgramCL(K.ADD_SCALAR).h.value != pcl.ProgramCL(K.ADD_SCALAR, "-DX=1").h.value


def test_compile_error_carries_the_log():
    with pytest.raises(ChaiError, match="undeclared identifier 'nothing_here'"):
        pcl.ProgramCL("__kernel void bad(__global float* a) { a[0] = nothing_here; }")


def test_unknown_kernel_and_missing_options():
    with pytest.raises(ChaiError, match="no kernel named"):
        pcl.ProgramCL(K.ADD_SCALAR).args("nope")
    with pytest.raises(ChaiError, match="STEPS"):
        pcl.ProgramCL(K.AMERICAN)            # STEPS / GROUP undefined


def test_local_element_sizes():
    assert pcl.element_size("float*") == 4
    assert pcl.element_size("double*") == 8
    assert pcl.element_size("float4*") == 16
    assert pcl.element_size("uchar*") == 1
    assert pcl.element_size("float3*") == 16


def _check(prog, name, args, g, l):
    L = pcl._bind(pcl.lib())
    A = (pcl.ClArg * len(args))()
    for i, (k, p, v) in enumerate(args):
        A[i].kind, A[i].prec = k, p
        if k == pcl.LOCAL:
            A[i].v.count = v
        elif k == pcl.SCALAR and p in (F32, 8):
            A[i].v.d = v
        elif k == pcl.SCALAR:
            A[i].v.u = v
    G = (ctypes.c_uint64 * 2)(*g)
    Lc = (ctypes.c_uint64 * 2)(*l)
    rc = L.chai_hip_cl_check(prog.kernel_handle(name), A, len(args), 1 if g[1] == 0 else 2, G, Lc)
    return rc, L.chai_hip_last_error().decode()


def test_launch_checks_before_running():
    p = pcl.ProgramCL(K.ADD_SCALAR)
    assert _check(p, "add_scalar", [(pcl.ARRAY, 0, 0), (pcl.SCALAR, F32, 1.5)], (20, 0), (1, 0))[0] == 0
    rc, msg = _check(p, "add_scalar", [(pcl.ARRAY, 0, 0)], (20, 0), (1, 0))
    assert rc != 0 and "takes 2 arguments" in msg
    rc, msg = _check(p, "add_scalar", [(pcl.ARRAY, 0, 0), (pcl.SCALAR, 8, 1.5)], (20, 0), (1, 0))
    assert rc != 0 and "4 bytes" in msg
    rc, msg = _check(p, "add_scalar", [(pcl.SCALAR, F32, 1.0), (pcl.SCALAR, F32, 1.5)], (20, 0), (1, 0))
    assert rc != 0 and "global_buffer" in msg
    rc, msg = _check(p, "add_scalar", [(pcl.ARRAY, 0, 0), (pcl.SCALAR, F32, 1.5)], (20, 0), (8, 0))
    assert rc != 0 and "not a multiple" in msg
    rc, msg = _check(p, "add_scalar", [(pcl.ARRAY, 0, 0), (pcl.SCALAR, F32, 1.5)], (4096, 0), (2048, 0))
    assert rc != 0 and "exceeds" in msg


def test_local_argument_wrapper_compiles():
    p = pcl.ProgramCL(K.BLOCK_SUM)
    args = [(pcl.ARRAY, 0, 0), (pcl.ARRAY, 0, 0), (pcl.LOCAL, 0, 64 * 4), (pcl.SCALAR, U32, 100)]
    rc, msg = _check(p, "block_sum", args, (128, 0), (64, 0))
    assert rc == 0, msg
    args[2] = (pcl.LOCAL, 0, 0)
    rc, msg = _check(p, "block_sum", args, (128, 0), (64, 0))
    assert rc != 0 and "local argument" in msg


def _inline_trace(n=20):
    # C = A + B; (progCL, "add_scalar", C, 1.5f)(N, 1); C += A - 3.3f  (sample/inline_cl.cpp)
    p = pcl.ProgramCL(K.ADD_SCALAR)
    t = tr.Trace()
    A, B = t.input(F32, n), t.input(F32, n)
    C = t.output(F32, n)
    t.assign(C, A + B)
    p("add_scalar", C, 1.5)(n, 1)
    t.assign(C, C + (A - tr.f32(3.3)))
    return t, (A, B, C)


def test_trace_with_inline_kernel_compiles_into_three_kernels():
    t, _ = _inline_trace()
    assert "trace_vm" in tr.emit_source(t, 0) or len(tr.emit_source(t, 0)) > 0
    assert len(tr.emit_source(t, 2)) > 0
    with pytest.raises(ChaiError, match="not a fused VM kernel"):
        tr.emit_source(t, 1)
    with pytest.raises(ChaiError):
        tr.emit_source(t, 3)


def test_trace_inline_kernel_is_checked_at_compile_time():
    p = pcl.ProgramCL(K.ADD_SCALAR)
    t = tr.Trace()
    C = t.output(F32, 20)
    t.assign(C, tr.f32(1.0) + t.index(F32, 0, 20))
    p("add_scalar", C, 1.5)(20, 8)                 # 20 is not a multiple of 8
    with pytest.raises(ChaiError, match="not a multiple"):
        tr.emit_source(t, 0)


def test_trace_inline_kernel_needs_its_first_array_as_lhs():
    p = pcl.ProgramCL(K.ADD_SCALAR)
    t = tr.Trace()
    C = t.output(F32, 20)
    t.assign(C, tr.f32(1.0) + t.index(F32, 0, 20))
    p("add_scalar", C, 1.5)(20, 1)
    lhs, nodes = t.stmts[-1]
    t.stmts[-1] = (lhs + 1, nodes)                 # a statement that claims another variable
    t.vars.append((F32, tr.TEMP, 20, 1, 1))
    with pytest.raises(ChaiError, match="first array argument"):
        tr.emit_source(t, 0)


# two independent scalar multiply-adds on neighbouring elements: at -O3 the
# gfx950 back end pairs such float code into packed FP32 instructions unless
# the packed-fp32-ops target feature is off
PAIRED_FMA = """
__kernel void axpb2(__global float* a, float s, float t)
{
    size_t i = get_global_id(0);
    a[2 * i] = a[2 * i] * s + t;
    a[2 * i + 1] = a[2 * i + 1] * s + t;
}
"""
OBJDUMP = "/opt/rocm/llvm/bin/llvm-objdump"


@pytest.mark.skipif(not os.path.exists(OBJDUMP), reason="llvm-objdump not installed")
def test_inline_kernels_carry_no_packed_fp32_instructions():
    code = pcl.ProgramCL(PAIRED_FMA).code_object()
    assert code[:4] == b"\x7fELF"
    with tempfile.NamedTemporaryFile(suffix=".co") as f:
        f.write(code)
        f.flush()
        asm = subprocess.run([OBJDUMP, "-d", "--mcpu=gfx950", f.name], capture_output=True, text=True,
                             check=True).stdout
    assert re.search(r"\bv_(fma|mul|add)_f32", asm), "kernel body not found in the disassembly"
    assert not re.search(r"\bv_pk_(fma|mul|add)_f32", asm), asm
