"""ProgramCL: inline OpenCL C kernels on the back end's arrays.

Mirrors the reference's ProgramCL (api/chai.h:703-771, api/ProgramCL.cpp):

    prog = ProgramCL("__kernel void add_scalar(__global float* a, float s)"
                     "{ a[get_global_id(0)] += s; }")
    prog("add_scalar", C, 1.5)(N, 1)          # (progCL, "add_scalar", C, 1.5f)(N, 1)

Arguments follow the kernel's parameter list: an array (a CUDA tensor, or a
trace variable) for each __global pointer, a number for each scalar
parameter (converted to the parameter's declared type: uint, int, float or
double), and an element count for each __local pointer (as the reference's
OCLLocalbuf<T>(length), memmgr/MemTransJIT.cpp:209-230).  The work space is
(global0, local0) or (global0, global1, local0, local1), in work items.

With CUDA tensors the kernel is launched at once on the current stream.  With
trace variables the call is recorded in their trace as the statement the
reference records (kernel_from_opencl, api/ProgramCL.cpp:202-219): it runs in
order between the fused kernels of the trace, on the trace's buffers.

One deliberate difference: the reference stores a C++ argument in the union
member of its own C++ type and reads the member of the parameter's type
(jitast/AstOpenCL.cpp:89-181), so a size_t passed for a uint parameter
arrives as the low word of a double.  Here the value is converted to the
parameter's type.
"""
import ctypes
import re

from . import F32, F64, I32, U32, ChaiError, check, init, lib, _stream

ARRAY, LOCAL, SCALAR, OTHER = 0, 1, 2, 3
_SCALAR = {"uint": U32, "unsigned int": U32, "unsigned": U32, "int": I32, "float": F32, "double": F64}
_BASE_SIZE = {"char": 1, "uchar": 1, "short": 2, "ushort": 2, "half": 2, "int": 4, "uint": 4, "float": 4,
              "long": 8, "ulong": 8, "double": 8}


class _ArgValue(ctypes.Union):
    _fields_ = [("ptr", ctypes.c_void_p), ("count", ctypes.c_uint64), ("u", ctypes.c_uint32),
                ("i", ctypes.c_int32), ("f", ctypes.c_float), ("d", ctypes.c_double)]


class ClArg(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("prec", ctypes.c_uint32), ("v", _ArgValue)]


class ClArgInfo(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("size", ctypes.c_uint32), ("type_name", ctypes.c_char * 64),
                ("name", ctypes.c_char * 64)]


def _bind(L):
    vp, sz, u32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32
    u64p = ctypes.POINTER(ctypes.c_uint64)
    L.chai_hip_cl_build.argtypes = [ctypes.c_char_p, sz, ctypes.c_char_p, ctypes.POINTER(vp)]
    L.chai_hip_cl_kernel_args.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ClArgInfo), ctypes.c_int]
    L.chai_hip_cl_kernel.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(vp)]
    L.chai_hip_cl_launch.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ClArg), u32, u32, u64p, u64p, vp]
    L.chai_hip_cl_check.argtypes = [vp, ctypes.POINTER(ClArg), u32, u32, u64p, u64p]
    L.chai_hip_cl_code_object.argtypes = [vp, vp, sz, ctypes.POINTER(sz)]
    return L


def element_size(type_name):
    """bytes of the pointee of an OpenCL pointer type name ("float*", "float4*")"""
    m = re.match(r"\s*(?:const\s+)?([a-z]+?)(\d*)\s*\*", type_name)
    if not m or m.group(1) not in _BASE_SIZE:
        raise ChaiError("cannot size __local element type %r" % type_name)
    n = int(m.group(2) or 1)
    return _BASE_SIZE[m.group(1)] * (4 if n == 3 else n)


class ProgramCL:
    """OpenCL C program text (a string, or a list of strings joined as
    clCreateProgramWithSource joins them), compiled once for gfx950 and
    cached by its text."""

    def __init__(self, text, options=""):
        if not isinstance(text, str):
            text = "".join(text)
        self.text = text
        self.options = options
        L = _bind(lib())
        h = ctypes.c_void_p()
        raw = text.encode()
        check(L.chai_hip_cl_build(raw, len(raw), options.encode(), ctypes.byref(h)), "ProgramCL build")
        self.h = h
        self._args = {}
        self._kernels = {}

    def args(self, name):
        """[(kind, size, type_name, name)] of kernel `name`'s parameters"""
        if name not in self._args:
            L = _bind(lib())
            n = ctypes.c_int()
            info = (ClArgInfo * 64)()
            check(L.chai_hip_cl_kernel_args(self.h, name.encode(), ctypes.byref(n), info, 64), "ProgramCL args")
            self._args[name] = [(a.kind, a.size, a.type_name.decode(), a.name.decode()) for a in info[:n.value]]
        return self._args[name]

    def code_object(self):
        """the gfx950 executable code object the text compiled into (bytes)"""
        L = _bind(lib())
        n = ctypes.c_size_t()
        check(L.chai_hip_cl_code_object(self.h, None, 0, ctypes.byref(n)), "ProgramCL code object")
        buf = ctypes.create_string_buffer(n.value)
        check(L.chai_hip_cl_code_object(self.h, buf, n.value, ctypes.byref(n)), "ProgramCL code object")
        return buf.raw[:n.value]

    def kernel_handle(self, name):
        if name not in self._kernels:
            L = _bind(lib())
            k = ctypes.c_void_p()
            check(L.chai_hip_cl_kernel(self.h, name.encode(), ctypes.byref(k)), "ProgramCL kernel")
            self._kernels[name] = k.value
        return self._kernels[name]

    def __call__(self, name, *args):
        """(progCL, "name", args...) -- call the result with the work space"""
        return KernelCall(self, name, args)


def _scalar_prec(type_name, size):
    t = type_name.replace("const ", "").strip()
    if t in _SCALAR:
        return _SCALAR[t]
    raise ChaiError("scalar parameter type %r (%d bytes) is not uint, int, float or double" % (type_name, size))


class KernelCall:
    def __init__(self, prog, name, args):
        self.prog, self.name, self.args = prog, name, list(args)

    def _work(self, dims):
        if len(dims) == 2:
            return 1, (int(dims[0]), 0), (int(dims[1]), 0)
        if len(dims) == 4:
            return 2, (int(dims[0]), int(dims[1])), (int(dims[2]), int(dims[3]))
        raise ChaiError("work space is (global0, local0) or (global0, global1, local0, local1)")

    def _values(self):
        """(kind, prec, value) per argument, converted to the parameter types"""
        params = self.prog.args(self.name)
        if len(params) != len(self.args):
            raise ChaiError("kernel %s takes %d arguments, %d given" % (self.name, len(params), len(self.args)))
        out = []
        for (kind, size, tname, pname), a in zip(params, self.args):
            if kind == ARRAY:
                out.append((ARRAY, 0, a))
            elif kind == LOCAL:
                out.append((LOCAL, 0, int(a) * element_size(tname)))
            elif kind == SCALAR:
                p = _scalar_prec(tname, size)
                v = float(a) if p in (F32, F64) else int(a)
                out.append((SCALAR, p, v))
            else:
                raise ChaiError("parameter %s of %s has an unsupported kind" % (pname, self.name))
        return out

    def __call__(self, *dims):
        ndim, g, l = self._work(dims)
        vals = self._values()
        from .trace import Var
        vars_ = [v for k, _, v in vals if k == ARRAY and isinstance(v, Var)]
        if vars_:
            tr = vars_[0].trace
            if any(v.trace is not tr for v in vars_) or len(vars_) != sum(k == ARRAY for k, _, _ in vals):
                raise ChaiError("array arguments of one inline kernel must all be variables of one trace")
            tr.opencl(self.prog.kernel_handle(self.name), vals, ndim, g, l)
            return
        init()
        L = _bind(lib())
        A = (ClArg * max(len(vals), 1))()
        for i, (k, p, v) in enumerate(vals):
            A[i].kind, A[i].prec = k, p
            if k == ARRAY:
                A[i].v.ptr = v.data_ptr()
            elif k == LOCAL:
                A[i].v.count = v
            elif p == U32:
                A[i].v.u = v & 0xffffffff
            elif p == I32:
                A[i].v.i = v
            elif p == F32:
                A[i].v.f = v
            else:
                A[i].v.d = v
        G = (ctypes.c_uint64 * 2)(*g)
        Lc = (ctypes.c_uint64 * 2)(*l)
        check(L.chai_hip_cl_launch(self.prog.h, self.name.encode(), A, len(vals), ndim, G, Lc, _stream()),
              "ProgramCL launch %s" % self.name)
