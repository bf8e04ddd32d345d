"""Trace builder: Chai-style array expressions recorded as the postfix IR of
include/chai_mi355x.h (chai_trace), the form a VectorTrace's bytecode takes
when it reaches the device (bytecode/ByteCodes.hpp opcodes, postfix order).

    t = Trace()
    z = t.rng_normal(F32, W=1 << 24, seed=7)           # rng_normal_make(gen, W)
    s = 100.0 * exp(f32(0.03) + f32(0.2) * z) ...      # Chai promotion rules
    m = t.assign(t.output(F32, 1), mean(payoff))
    prog = t.compile(); prog.run({m: out_tensor})

Scalar literal typing follows api/ArrayLit.cpp: a Python float is an f64
literal (promotes an f32 expression to f64, as a C++ double literal does in
Chai); wrap with f32() for a float literal; a Python int is an i32 literal.
"""
import ctypes

from . import F32, F64, I32, U32, check, init, lib, _stream

VAR = 0x10000
INPUT, OUTPUT, TEMP = 0, 1, 2


class TraceVar(ctypes.Structure):
    _fields_ = [("prec", ctypes.c_uint32), ("role", ctypes.c_uint32), ("W", ctypes.c_uint64),
                ("H", ctypes.c_uint64), ("slots", ctypes.c_uint64)]


class TraceNode(ctypes.Structure):
    _fields_ = [("code", ctypes.c_uint32), ("arg", ctypes.c_uint32), ("vec_len", ctypes.c_uint32),
                ("flags", ctypes.c_uint32), ("s0", ctypes.c_double), ("s1", ctypes.c_double),
                ("seed", ctypes.c_uint64), ("W", ctypes.c_uint64), ("H", ctypes.c_uint64),
                ("slots", ctypes.c_uint64)]


class TraceStmt(ctypes.Structure):
    _fields_ = [("lhs", ctypes.c_uint32), ("first", ctypes.c_uint32), ("count", ctypes.c_uint32),
                ("pad", ctypes.c_uint32)]


class TraceC(ctypes.Structure):
    _fields_ = [("vars", ctypes.POINTER(TraceVar)), ("num_vars", ctypes.c_uint32),
                ("stmts", ctypes.POINTER(TraceStmt)), ("num_stmts", ctypes.c_uint32),
                ("nodes", ctypes.POINTER(TraceNode)), ("num_nodes", ctypes.c_uint32)]


# ---------------------------------------------------------------- expressions
class Expr:
    def __init__(self, code, kids=(), **node):
        self.code = code
        self.kids = list(kids)
        self.node = node

    # arithmetic (bytecode/ByteCodes.hpp operators)
    def __add__(self, o): return Expr(11, [self, lit(o)])
    def __radd__(self, o): return Expr(11, [lit(o), self])
    def __sub__(self, o): return Expr(14, [self, lit(o)])
    def __rsub__(self, o): return Expr(14, [lit(o), self])
    def __mul__(self, o): return Expr(13, [self, lit(o)])
    def __rmul__(self, o): return Expr(13, [lit(o), self])
    def __truediv__(self, o): return Expr(12, [self, lit(o)])
    def __rtruediv__(self, o): return Expr(12, [lit(o), self])
    def __mod__(self, o): return Expr(17, [self, lit(o)])
    def __lshift__(self, o): return Expr(18, [self, lit(o)])
    def __rshift__(self, o): return Expr(19, [self, lit(o)])
    def __and__(self, o): return Expr(20, [self, lit(o)])
    def __or__(self, o): return Expr(21, [self, lit(o)])
    def __xor__(self, o): return Expr(22, [self, lit(o)])
    def __neg__(self): return Expr(10, [self])
    def __invert__(self): return Expr(15, [self])
    def __eq__(self, o): return Expr(32, [self, lit(o)])
    def __ne__(self, o): return Expr(33, [self, lit(o)])
    def __ge__(self, o): return Expr(34, [self, lit(o)])
    def __gt__(self, o): return Expr(35, [self, lit(o)])
    def __le__(self, o): return Expr(36, [self, lit(o)])
    def __lt__(self, o): return Expr(37, [self, lit(o)])
    __hash__ = object.__hash__

    @property
    def T(self):
        """transpose of a matmul operand (folded into the GEMM, MakeStmt.cpp:704-720)"""
        e = Expr(self.code, self.kids, **dict(self.node))
        e.node["flags"] = e.node.get("flags", 0) ^ 1
        return e


class Var(Expr):
    def __init__(self, index):
        super().__init__(VAR, arg=index)
        self.index = index


def lit(v):
    if isinstance(v, Expr):
        return v
    if isinstance(v, bool):
        v = int(v)
    if isinstance(v, int):
        return Expr(6, s0=float(v))           # scalar_i32
    return Expr(8, s0=float(v))               # scalar_f64


def f32(v): return Expr(7, s0=float(v))
def f64(v): return Expr(8, s0=float(v))
def i32(v): return Expr(6, s0=float(v))
def u32(v): return Expr(5, s0=float(v))


def _u(code):
    return lambda x: Expr(code, [lit(x)])


exp, exp2, log, log2, log10, sqrt, rsqrt = _u(219), _u(220), _u(237), _u(238), _u(239), _u(261), _u(256)
sin, cos, tan, tanh, fabs, floor, ceil = _u(257), _u(214), _u(262), _u(263), _u(223), _u(225), _u(212)
rint, round_, trunc, erf, erfc, expm1, log1p = _u(253), _u(255), _u(266), _u(218), _u(217), _u(222), _u(240)
iabs, clz, lnot = _u(400), _u(404), _u(16)
sum_, mean = _u(132), _u(131)
asin, acos, atan, sinh, cosh, cbrt, exp10 = _u(203), _u(200), _u(206), _u(259), _u(215), _u(211), _u(221)
acosh, acospi, asinh, asinpi, atanh, atanpi = _u(201), _u(202), _u(204), _u(205), _u(208), _u(209)
cospi, sinpi, tanpi, ilogb, lgamma, logb, tgamma = _u(216), _u(260), _u(264), _u(233), _u(235), _u(241), _u(265)
nan, degrees, radians, sign = _u(246), _u(301), _u(305), _u(308)
isfinite, isinf, isnan, isnormal, signbit = _u(507), _u(508), _u(509), _u(510), _u(513)
convert_u32, convert_i32, convert_f32, convert_f64 = _u(1), _u(2), _u(3), _u(4)


def _b(code):
    return lambda a, b: Expr(code, [lit(a), lit(b)])


maxe, mine, pow_, atan2, fmod, hypot = _b(302), _b(303), _b(248), _b(207), _b(229), _b(232)
land, lor, mul_hi, rotate, add_sat, sub_sat = _b(30), _b(31), _b(412), _b(414), _b(402), _b(415)
hadd, rhadd, abs_diff, mul24, fmaxe, fmine = _b(405), _b(413), _b(401), _b(411), _b(227), _b(228)
dot = _b(130)
copysign, fdim, atan2pi, ldexp, maxmag, minmag = _b(213), _b(224), _b(210), _b(234), _b(243), _b(244)
nextafter, pown, powr, remainder, rootn, step = _b(247), _b(249), _b(250), _b(251), _b(254), _b(306)
isequal, isnotequal, isgreater, isgreaterequal = _b(500), _b(501), _b(502), _b(503)
isless, islessequal, islessgreater, isordered, isunordered = _b(504), _b(505), _b(506), _b(511), _b(512)


def cond(c, a, b): return Expr(23, [lit(c), lit(a), lit(b)])
def clamp(a, lo, hi): return Expr(300, [lit(a), lit(lo), lit(hi)])
def fma(a, b, c): return Expr(226, [lit(a), lit(b), lit(c)])
def mad(a, b, c): return Expr(242, [lit(a), lit(b), lit(c)])
def mix(a, b, c): return Expr(304, [lit(a), lit(b), lit(c)])
def smoothstep(e0, e1, x): return Expr(307, [lit(e0), lit(e1), lit(x)])
def mad24(a, b, c): return Expr(406, [lit(a), lit(b), lit(c)])
def mad_hi(a, b, c): return Expr(407, [lit(a), lit(b), lit(c)])
def mad_sat(a, b, c): return Expr(408, [lit(a), lit(b), lit(c)])


def _gather_data(a):
    if not isinstance(a, Expr):
        raise TypeError("gather data must be an array expression")
    return a        # an expression is computed into a temp by the compiler first


def gather1_floor(a, i):
    """out(x, y) = a(floor(i(x, y)) wrapped into [0, W), y), shape and type of a
    (api/chai.h:158, interp/InterpGather.hpp:20-30): a negative subscript gets W
    added once, the result is reduced modulo W."""
    return Expr(40, [_gather_data(a), lit(i)])


def gather2_floor(a, i, j):
    """out(x, y) = a(floor(i(x, y)) wrapped into [0, W), floor(j(x, y)) wrapped
    into [0, H)) (api/chai.h:165, interp/InterpGather.hpp:34-48)."""
    return Expr(41, [_gather_data(a), lit(i), lit(j)])


def matmul(a, b): return Expr(120, [a, b])
def matmul_g(a, b, c, alpha=1.0, beta=1.0): return Expr(121, [a, b, c], s0=alpha, s1=beta)


def set_jit(on):
    """Fused kernels: True = hipRTC-compiled straight-line HIP, False = trace VM kernel."""
    L = lib()
    L.chai_hip_trace_set_jit.argtypes = [ctypes.c_int]
    check(L.chai_hip_trace_set_jit(int(on)), "trace_set_jit")


def emit_source(trace, kernel=0):
    """JIT source of a trace's fused kernel, produced by the front end alone (no GPU)."""
    c = trace.to_c()
    L = lib()
    L.chai_hip_trace_emit_source.argtypes = [ctypes.POINTER(TraceC), ctypes.c_int, ctypes.c_char_p,
                                             ctypes.c_size_t]
    buf = ctypes.create_string_buffer(1 << 20)
    check(L.chai_hip_trace_emit_source(ctypes.byref(c), kernel, buf, len(buf)), "trace_emit_source")
    return buf.value.decode()


# --------------------------------------------------------------------- trace
class Trace:
    def __init__(self):
        self.vars = []
        self.stmts = []      # (lhs, [nodes])

    def var(self, prec, W, H=1, slots=1, role=TEMP):
        self.vars.append((prec, role, W, H, slots))
        v = Var(len(self.vars) - 1)
        v.trace = self
        return v

    def input(self, prec, W, H=1, slots=1):
        return self.var(prec, W, H, slots, INPUT)

    def output(self, prec, W, H=1, slots=1):
        return self.var(prec, W, H, slots, OUTPUT)

    def temp(self, prec, W, H=1, slots=1):
        return self.var(prec, W, H, slots, TEMP)

    def assign(self, lhs, expr):
        nodes = []
        self._flatten(lit(expr), nodes)
        self.stmts.append((lhs.index, nodes))
        return lhs

    def opencl(self, kernel, vals, ndim, glob, loc):
        """Record an inline OpenCL kernel statement (kernel_from_opencl, bytecode 50):
        argument nodes in order, then the root with the kernel handle and the work
        space.  `vals` are (kind, prec, value) from program_cl.KernelCall; the
        statement's lhs is the first array argument (api/ProgramCL.cpp:202-219)."""
        nodes = []
        for kind, prec, v in vals:
            if kind == 0:
                nodes.append((VAR, {"arg": v.index}))
            elif kind == 1:
                nodes.append((0x10001, {"W": int(v)}))
            else:
                nodes.append(({U32: 5, I32: 6, F32: 7, F64: 8}[prec], {"s0": float(v)}))
        nodes.append((50, {"arg": len(vals), "seed": int(kernel), "vec_len": ndim, "W": glob[0], "H": glob[1],
                           "s0": float(loc[0]), "s1": float(loc[1])}))
        first = next(v for kind, _, v in vals if kind == 0)
        self.stmts.append((first.index, nodes))

    def _flatten(self, e, out):
        """postfix order (children, then the node), without recursion so that
        deep expressions (long traced loops) flatten too"""
        stack = [(e, False)]
        while stack:
            x, done = stack.pop()
            if done:
                out.append((x.code, dict(x.node)))
                continue
            stack.append((x, True))
            for k in reversed(x.kids):
                stack.append((k, False))

    # data creation leaves (api/RngFun.cpp, IndexFun.cpp)
    @staticmethod
    def rng_uniform(prec, W, H=1, slots=1, seed=0, variant=1230, vec_len=0, limits=None):
        code = {U32: 112, I32: 113, F32: 114, F64: 115}[prec]
        lo, hi = limits if limits is not None else (0.0, 0.0)
        return Expr(code, arg=variant, vec_len=vec_len, seed=seed, W=W, H=H, slots=slots, s0=lo, s1=hi)

    @staticmethod
    def rng_normal(prec, W, H=1, slots=1, seed=0, variant=1230, vec_len=0):
        code = {F32: 110, F64: 111}[prec]
        return Expr(code, arg=variant, vec_len=vec_len, seed=seed, W=W, H=H, slots=slots)

    @staticmethod
    def index(prec, dim, W, H=1, slots=1):
        code = (42 if H == 1 else 46) + {U32: 0, I32: 1, F32: 2, F64: 3}[prec]
        return Expr(code, arg=dim, W=W, H=H, slots=slots)

    @staticmethod
    def ones(prec, W, H=1):
        """ones1_* / ones2_* (bytecode/ByteCodes.hpp:90-97).  Note that the
        reference declares zeros2_f64 with the code of ones2_i32 (73), so
        ones(I32, W, H > 1) is an f64 zero array, as in the reference."""
        return Expr((68 if H == 1 else 72) + {U32: 0, I32: 1, F32: 2, F64: 3}[prec], W=W, H=H)

    @staticmethod
    def zeros(prec, W, H=1):
        """zeros1_* / zeros2_* (bytecode/ByteCodes.hpp:98-105; zeros2_f64 is 73)"""
        if H > 1 and prec == F64:
            return Expr(73, W=W, H=H)
        return Expr((76 if H == 1 else 80) + {U32: 0, I32: 1, F32: 2, F64: 3}[prec], W=W, H=H)

    def to_c(self):
        """ctypes chai_trace (arrays kept alive on the returned object)."""
        nv, ns = len(self.vars), len(self.stmts)
        nn = sum(len(n) for _, n in self.stmts)
        V = (TraceVar * nv)()
        for i, (p, r, W, H, S) in enumerate(self.vars):
            V[i] = TraceVar(p, r, W, H, S)
        N = (TraceNode * max(nn, 1))()
        S_ = (TraceStmt * ns)()
        k = 0
        for i, (lhs, nodes) in enumerate(self.stmts):
            S_[i] = TraceStmt(lhs, k, len(nodes), 0)
            for code, nd in nodes:
                N[k] = TraceNode(code, nd.get("arg", 0), nd.get("vec_len", 0), nd.get("flags", 0),
                                 nd.get("s0", 0.0), nd.get("s1", 0.0), nd.get("seed", 0),
                                 nd.get("W", 0), nd.get("H", 0), nd.get("slots", 0))
                k += 1
        c = TraceC(V, nv, S_, ns, N, nn)
        c._keep = (V, N, S_)
        return c

    def compile(self):
        init()
        c = self.to_c()
        h = ctypes.c_void_p()
        L = lib()
        L.chai_hip_trace_compile.argtypes = [ctypes.POINTER(TraceC), ctypes.POINTER(ctypes.c_void_p)]
        check(L.chai_hip_trace_compile(ctypes.byref(c), ctypes.byref(h)), "chai_hip_trace_compile")
        return Program(self, h, c)


def cache_stats():
    """(hits, misses, entries) of the compiled-trace cache."""
    init()
    L = lib()
    L.chai_hip_trace_cache_stats.argtypes = [ctypes.POINTER(ctypes.c_uint64)] * 3
    h, m, e = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
    check(L.chai_hip_trace_cache_stats(ctypes.byref(h), ctypes.byref(m), ctypes.byref(e)), "trace_cache_stats")
    return h.value, m.value, e.value


class Program:
    def __init__(self, trace, handle, c):
        self.trace = trace
        self.h = handle
        self._c = c

    def info(self):
        k, n = ctypes.c_int(), ctypes.c_int()
        lib().chai_hip_trace_info.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                                              ctypes.POINTER(ctypes.c_int)]
        check(lib().chai_hip_trace_info(self.h, ctypes.byref(k), ctypes.byref(n)), "trace_info")
        return k.value, n.value

    def source(self, kernel=0):
        """(generated JIT source or JIT log, jitted?) of fused kernel `kernel`."""
        L = lib()
        L.chai_hip_trace_kernel_source.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p,
                                                   ctypes.c_size_t, ctypes.POINTER(ctypes.c_int)]
        buf = ctypes.create_string_buffer(1 << 20)
        j = ctypes.c_int()
        check(L.chai_hip_trace_kernel_source(self.h, kernel, buf, len(buf), ctypes.byref(j)), "trace_kernel_source")
        return buf.value.decode(errors="replace"), bool(j.value)

    def run(self, buffers):
        """buffers: {Var or index: cuda tensor} for every INPUT and OUTPUT variable."""
        n = len(self.trace.vars)
        arr = (ctypes.c_void_p * n)()
        for v, t in buffers.items():
            idx = v.index if isinstance(v, Var) else int(v)
            arr[idx] = t.data_ptr()
        L = lib()
        L.chai_hip_trace_launch.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p]
        check(L.chai_hip_trace_launch(self.h, arr, _stream()), "chai_hip_trace_launch")

    def __del__(self):
        try:
            lib().chai_hip_trace_release.argtypes = [ctypes.c_void_p]
            lib().chai_hip_trace_release(self.h)
        except Exception:
            pass
