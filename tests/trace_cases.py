"""Trace workloads shared by the CPU oracle tests and the GPU parity tests."""
import hashlib
import math
import struct

import numpy as np

from chai_amd import F32, F64, I32, U32
from chai_amd.trace import (Trace, cond, exp, f32, fabs, fma, hadd, log, lor, maxe, mean, mine,
                            mul_hi, rotate, sqrt, sum_, u32, i32, add_sat, sub_sat, rhadd,
                            abs_diff, mul24, clamp, dot, gather1_floor, gather2_floor, matmul)

# ---------------------------------------------------------------- MD5 (RFC 1321)
MD5_S = [7, 12, 17, 22] * 4 + [5, 9, 14, 20] * 4 + [4, 11, 16, 23] * 4 + [6, 10, 15, 21] * 4
MD5_K = [int(abs(math.sin(i + 1)) * 2 ** 32) & 0xffffffff for i in range(64)]


def md5_words(msg):
    """one padded 64-byte block as 16 little-endian u32 words (len(msg) <= 55)"""
    assert len(msg) <= 55
    b = msg + b"\x80" + b"\x00" * (55 - len(msg)) + struct.pack("<Q", 8 * len(msg))
    return list(struct.unpack("<16I", b))


def md5_trace(n):
    """MD5 of n single-block messages as a pure uint32 trace (the reference's
    sample/md5.cpp workload): 16 input word arrays, 4 digest outputs."""
    t = Trace()
    M = [t.input(U32, n) for _ in range(16)]
    A, B, C, D = u32(0x67452301), u32(0xefcdab89), u32(0x98badcfe), u32(0x10325476)
    a0, b0, c0, d0 = A, B, C, D
    for i in range(64):
        if i < 16:
            F, g = (B & C) | (~B & D), i
        elif i < 32:
            F, g = (D & B) | (~D & C), (5 * i + 1) % 16
        elif i < 48:
            F, g = B ^ C ^ D, (3 * i + 5) % 16
        else:
            F, g = C ^ (B | ~D), (7 * i) % 16
        nb = t.assign(t.temp(U32, n), B + rotate(F + A + u32(MD5_K[i]) + M[g], u32(MD5_S[i])))
        A, D, C, B = D, C, B, nb
    outs = [t.assign(t.output(U32, n), x + y) for x, y in ((a0, A), (b0, B), (c0, C), (d0, D))]
    return t, M, outs


def md5_inputs(msgs):
    words = np.array([md5_words(m) for m in msgs], dtype=np.uint32)   # [n][16]
    return [np.ascontiguousarray(words[:, k]) for k in range(16)]


def md5_expected(msgs):
    return [hashlib.md5(m).hexdigest() for m in msgs]


def md5_hex(outs):
    n = len(outs[0])
    return [b"".join(struct.pack("<I", int(o[i])) for o in outs).hex() for i in range(n)]


MSGS = [("chai md5 message %d %s" % (i, "x" * (i % 23))).encode() for i in range(28)]

# ------------------------------------------------------------ Monte Carlo trace
from chai_amd.workloads import mc_trace  # noqa: E402,F401  (shared with bench.py)


# ------------------------------------------------------------- random traces
INT_BIN = ["add", "sub", "mul", "and", "or", "xor", "shl", "shr", "max", "min", "lt", "eq",
           "rotate", "add_sat", "sub_sat", "hadd", "rhadd", "abs_diff", "mul24", "mul_hi", "div", "mod"]
FP_BIN = ["add", "sub", "mul", "div", "max", "min", "lt", "ge", "eq"]


def _bin(name, a, b):
    import operator as op
    table = {"add": op.add, "sub": op.sub, "mul": op.mul, "div": op.truediv, "mod": op.mod,
             "and": op.and_, "or": op.or_, "xor": op.xor, "shl": op.lshift, "shr": op.rshift,
             "max": maxe, "min": mine, "lt": op.lt, "ge": op.ge, "eq": op.eq, "rotate": rotate,
             "add_sat": add_sat, "sub_sat": sub_sat, "hadd": hadd, "rhadd": rhadd,
             "abs_diff": abs_diff, "mul24": mul24, "mul_hi": mul_hi}
    return table[name](a, b)


def random_trace(seed, kind="int", n_in=3, W=37, H=5, slots=2, depth=4, n_stmt=3):
    """A random multi-statement trace with broadcast operands (smaller inputs),
    temps, an output per statement and a sum reduction.  kind: int | fp."""
    rng = np.random.default_rng(seed)
    t = Trace()
    precs = [U32, I32] if kind == "int" else [F32, F64, I32]
    ins, data = [], {}
    for k in range(n_in):
        p = precs[rng.integers(len(precs))]
        w = W if rng.random() < 0.7 else 1 + int(rng.integers(W))
        h = H if rng.random() < 0.7 else 1
        s = slots if rng.random() < 0.7 else 1
        v = t.input(p, w, h, s)
        ins.append(v)
        if p == U32:
            data[v.index] = rng.integers(0, 2 ** 32, size=(s, h, w), dtype=np.uint64).astype(np.uint32)
        elif p == I32:
            lo, hi = (-2 ** 31, 2 ** 31) if kind == "int" else (-50, 50)
            data[v.index] = rng.integers(lo, hi, size=(s, h, w), dtype=np.int64).astype(np.int32)
        else:
            data[v.index] = (rng.random((s, h, w)) * 4 - 2).astype(np.float32 if p == F32 else np.float64)
    ops = INT_BIN if kind == "int" else FP_BIN

    def gen(d, pool):
        if d == 0 or rng.random() < 0.25:
            return pool[rng.integers(len(pool))]
        name = ops[rng.integers(len(ops))]
        a, b = gen(d - 1, pool), gen(d - 1, pool)
        if kind == "fp" and name == "div":
            b = fabs(b + f32(0.0)) + f32(0.5)      # (+ 0.0f lifts an int operand to f32)
        return _bin(name, a, b)

    pool = list(ins)
    outs = []
    for i in range(n_stmt):
        e = gen(depth, pool)
        tv = t.temp(F64 if kind == "fp" else U32, W, H, slots) if i < n_stmt - 1 else None
        ov = t.output(F64 if kind == "fp" else U32, W, H, slots)
        # full-shape anchor so the statement has the index space of its variable
        anchor = u32(0) if kind == "int" else f32(0.0)
        full = t.rng_uniform(U32 if kind == "int" else F32, W, H, slots, seed=seed + i)
        e = e + (full & u32(0)) if kind == "int" else e + full * f32(0.0)
        t.assign(ov, e)
        outs.append(ov)
        if tv is not None:
            t.assign(tv, e * anchor + e)
            pool.append(tv)
    red = t.assign(t.output(F64 if kind == "fp" else U32, 1, 1, slots), sum_(outs[-1]))
    return t, data, outs, red


def bcast_trace(seed, W=37, H=5, slots=2, depth=4, n_stmt=3):
    """Nested modulo broadcast: integer leaves (inputs, rng and index arrays) of
    many small shapes combined so that intermediate results are themselves
    smaller than the statement's index space (interp/BaseInterp.cpp:50-55
    evaluates every operator into an array of its own dims)."""
    rng = np.random.default_rng(10_000 + seed)
    t = Trace()
    widths = [1, 2, 3, 5, 6, 12, 13, 24, W]
    heights = [1, H]
    sl = [1, slots]
    data, leaves = {}, []
    for k in range(4):
        p = [U32, I32][rng.integers(2)]
        w, h, s = int(rng.choice(widths)), int(rng.choice(heights)), int(rng.choice(sl))
        v = t.input(p, w, h, s)
        data[v.index] = rng.integers(-1000, 1000, size=(s, h, w), dtype=np.int64).astype(
            np.int32 if p == I32 else np.uint32)
        leaves.append(lambda v=v: v)
    for k in range(3):
        w, h, s = int(rng.choice(widths)), int(rng.choice(heights)), int(rng.choice(sl))
        sd = int(rng.integers(1000))
        leaves.append(lambda w=w, h=h, s=s, sd=sd: Trace.rng_uniform(U32, w, h, s, seed=sd))
        w2, h2, s2 = int(rng.choice(widths)), int(rng.choice(heights)), int(rng.choice(sl))
        dim = int(rng.integers(2))
        leaves.append(lambda w=w2, h=h2, s=s2, dim=dim: Trace.index(U32, dim, w, h, s))
    ops = ["add", "sub", "mul", "xor", "max", "min", "and", "or"]

    def gen(d):
        if d == 0 or rng.random() < 0.3:
            return leaves[rng.integers(len(leaves))]()
        return _bin(ops[rng.integers(len(ops))], gen(d - 1), gen(d - 1))

    outs = []
    for i in range(n_stmt):
        e = gen(depth)
        full = t.rng_uniform(U32, W, H, slots, seed=500 + i)
        outs.append(t.assign(t.output(U32, W, H, slots), e + (full & u32(0))))
    return t, data, outs


# ------------------------------------------------------------------- gathers
from chai_amd.workloads import filter_expected, filter_trace  # noqa: E402,F401  (shared with bench.py)


LATTICE = dict(up=1.0207, pr_up=0.5102, pr_down=0.4898, disc=1.0005, strike=100.0)


def lattice_trace(N, steps, up, pr_up, pr_down, disc, strike):
    """sample/finance/american_option.cpp's backward induction, `steps` steps
    in one trace: each step gathers the previous step's option values at IDX
    and IDX + 1, a temp computed in the same trace, so every step is a kernel
    of its own."""
    t = Trace()
    prices, option = t.input(F64, N), t.input(F64, N)
    ins = (prices, option)
    IDX = Trace.index(F64, 0, N)
    for k in range(steps):
        mk = t.output if k == steps - 1 else t.temp
        no_ex = (pr_down * gather1_floor(option, IDX) + pr_up * gather1_floor(option, IDX + 1)) / disc
        prices = t.assign(mk(F64, N), up * prices)
        option = t.assign(mk(F64, N), maxe(no_ex, prices - strike))
    return t, ins, (prices, option)


def lattice_expected(prices, option, steps, up, pr_up, pr_down, disc, strike):
    for _ in range(steps):
        no_ex = (pr_down * option + pr_up * np.roll(option, -1)) / disc
        prices = up * prices
        ex = prices - strike
        option = np.where(no_ex < ex, ex, no_ex)       # max(a, b) = a < b ? b : a
    return prices, option


def gather_subscripts(prec, shape, span, rng):
    """subscript arrays of one type covering every range of the wrap rule:
    in [0, span), negative above and below -span, past span, and the
    conversion's special values"""
    n = int(np.prod(shape))
    if prec in (U32, I32):
        v = rng.integers(-3 * span - 2, 3 * span + 3, size=n).astype(np.int64)
        sp = [-2 ** 31, 2 ** 31 - 1, -2 ** 31 + 7, -span, -span - 1, span, 0, -1]
        v[:min(n, len(sp))] = sp[:n]
        return (v.astype(np.int32).view(np.uint32) if prec == U32 else v.astype(np.int32)).reshape(shape)
    v = rng.uniform(-3 * span - 2, 3 * span + 3, size=n)
    sp = [np.nan, np.inf, -np.inf, 3e9, -3e9, -0.0, -0.5, 1e-30, -1e-30, float(span) - 0.25,
          -float(span), 2.0 ** 31, -2.0 ** 31, float(-span) - 0.5]
    v[:min(n, len(sp))] = sp[:n]
    return v.astype(np.float32 if prec == F32 else np.float64).reshape(shape)


def subscript_closed_form(v, prec, W):
    """the wrap rule per range of the floored subscript xf (C's int / size_t
    conversions): xf >= 0 -> xf mod W; -W <= xf < 0 -> xf + W (0 at -W);
    xf < -W -> the int xf + W sign-extended to 64 bits, (2^64 + xf + W) mod W"""
    if prec == U32:
        xf = int(np.uint32(v).view(np.int32))
    elif prec == I32:
        xf = int(v)
    else:
        f = math.floor(float(v)) if math.isfinite(float(v)) else None
        xf = f if f is not None and -2 ** 31 <= f < 2 ** 31 else -2 ** 31
    if xf >= 0:
        return xf % W
    if xf >= -W:
        return (xf + W) % W
    return (2 ** 64 + xf + W) % W


def gather_rule_trace(prec_i, two, W=12, H=10, S=2, Si=3, seed=0):
    """data element (s, y, x) holds its own flat index; gathered through
    subscripts of type prec_i with more slots than the data (the output keeps
    the data's slots, the subscripts are read at slot % Si)."""
    rng = np.random.default_rng(seed)
    t = Trace()
    a = t.input(I32, W, H, S)
    i = t.input(prec_i, W, H, Si)
    j = t.input(prec_i, W, 2, 1) if two else None
    o = t.assign(t.output(I32, W, H, S), gather2_floor(a, i, j) if two else gather1_floor(a, i))
    data = {a.index: np.arange(S * H * W, dtype=np.int32).reshape(S, H, W),
            i.index: gather_subscripts(prec_i, (Si, H, W), W, rng)}
    if two:
        data[j.index] = gather_subscripts(prec_i, (1, 2, W), H, rng)
    expect = np.zeros((S, H, W), dtype=np.int32)
    for s_ in range(S):
        for y in range(H):
            for x in range(W):
                xr = subscript_closed_form(data[i.index][s_ % Si, y, x], prec_i, W)
                yr = subscript_closed_form(data[j.index][0, y % 2, x], prec_i, H) if two else y
                expect[s_, y, x] = (s_ * H + yr) * W + xr
    return t, data, o, expect


def gather_bcast_trace(seed=0):
    """gathers whose results are smaller than the statement (broadcast by
    modulo), a subscript expression mixing an input and an index array, a
    gather under a sum reduction, and a gather from a variable computed in
    the same trace (a second kernel)."""
    rng = np.random.default_rng(seed)
    t = Trace()
    a = t.input(F32, 5, 3, 2)
    i = t.input(I32, 5, 1, 1)
    b = t.input(F32, 10, 6, 4)
    yi = Trace.index(I32, 1, 5, 3)
    o1 = t.assign(t.output(F32, 10, 6, 4), gather1_floor(a, i) * b + gather2_floor(a, i * 3 - 7, yi + i))
    s1 = t.assign(t.output(F32, 1, 1, 2), sum_(gather1_floor(a, i + 1)))
    tmp = t.assign(t.temp(F32, 5, 3, 2), a * f32(2.0) + f32(1.0))
    o2 = t.assign(t.output(F32, 5, 3, 2), gather2_floor(tmp, Trace.index(F32, 0, 5) * f32(-1.5), yi - 1))
    A = rng.standard_normal((2, 3, 5)).astype(np.float32)
    I = rng.integers(-20, 20, size=(1, 1, 5)).astype(np.int32)
    B = rng.standard_normal((4, 6, 10)).astype(np.float32)
    wrap = lambda v, W: subscript_closed_form(v, I32, W)  # noqa: E731
    g1 = A[:, :, [wrap(v, 5) for v in I.ravel()]]                       # (2, 3, 5)
    g2 = np.empty_like(A)
    for y in range(3):
        for x in range(5):
            g2[:, y, x] = A[:, wrap(y + I[0, 0, x], 3), wrap(I[0, 0, x] * 3 - 7, 5)]
    tile = lambda g: g[np.arange(4)[:, None, None] % g.shape[0], np.arange(6)[None, :, None] % g.shape[1],  # noqa: E731
                       np.arange(10)[None, None, :] % g.shape[2]]
    e1 = tile(g1) * B + tile(g2)
    g3 = A[:, :, [wrap(v + 1, 5) for v in I.ravel()]]
    e_s1 = g3.astype(np.float64).reshape(2, -1).sum(1).astype(np.float32)   # exact in double
    T = A * np.float32(2.0) + np.float32(1.0)
    e2 = np.empty_like(A)
    for y in range(3):
        for x in range(5):
            xf = np.float32(x) * np.float32(-1.5)
            e2[:, y, x] = T[:, wrap(y - 1, 3), subscript_closed_form(xf, F32, 5)]
    return t, {a.index: A, i.index: I, b.index: B}, (o1, s1, o2), (e1, e_s1, e2)


def kirchhoff_trace(NT, N, NNN=12):
    """The reference's sample/kirch.cpp (Kirchhoff migration): modl starts as
    zeros_f32 and accumulates N gather2_floor reads of data at (it, index1),
    so modl is assigned N + 1 times in one trace.  Literal types follow the
    C++: float variables are f32 literals, 0.5 is a double."""
    f = np.float32
    dx, dt = f(f(NNN) / f(N)), f(f(NNN) / f(NT))
    factor, idt = f(1.0 / (f(1) * f(1))), f(1.0 / float(dt))
    t = Trace()
    data = t.input(F32, NT, N)
    modl = t.output(F32, NT, N)
    t.assign(modl, Trace.zeros(F32, NT, N))
    x = t.assign(t.temp(F32, NT, N), f32(dx) * Trace.index(F32, 1, NT, N))
    z = t.assign(t.temp(F32, NT, N), f32(dt) * Trace.index(F32, 0, NT, N))
    for iy in range(N):
        y = f(f(iy) * dx)
        index1 = f32(float(iy)) * Trace.ones(F32, NT, N)
        it = 0.5 + sqrt(z * z + (x - f32(y)) * (x - f32(y)) * f32(factor)) * f32(idt)
        t.assign(modl, modl + gather2_floor(data, it, index1))
    return t, data, modl, (dx, dt, factor, idt)


def kirchhoff_expected(d, NT, N, consts):
    """numpy restatement of sample/kirch.cpp on data [N][NT] (H = N rows of W = NT)"""
    f = np.float32
    dx, dt, factor, idt = consts
    yy, xx = np.meshgrid(np.arange(N), np.arange(NT), indexing="ij")
    X, Z = f(dx) * yy.astype(f), f(dt) * xx.astype(f)
    modl = np.zeros((N, NT), dtype=f)
    for iy in range(N):
        y = f(f(iy) * dx)
        s = Z * Z + (X - y) * (X - y) * factor
        it = 0.5 + (np.sqrt(s) * idt).astype(np.float64)
        xs = np.array([subscript_closed_form(v, F64, NT) for v in it.ravel()]).reshape(N, NT)
        modl = modl + d[iy % N, xs]
    return modl



def loopsum_trace():
    """sample/loopsum.cpp (one instance): B = 5 * ones_f32(1); twice { three
    times A = A + mean(B); B += 1 }; C = sum(A).  mean sits inside the
    statement; the compiler splits it out."""
    t = Trace()
    A = t.input(F32, 100)
    An = t.output(F32, 100)
    B = t.temp(F32, 1)
    C = t.output(F32, 1)
    t.assign(B, 5 * Trace.ones(F32, 1))
    cur = A
    for j in range(2):
        for k in range(3):
            cur = t.assign(An, cur + mean(B))
        t.assign(B, B + 1)
    t.assign(C, sum_(An))
    return t, A, C, An


def split_trace_case():
    """statements with a reduction, a matmul and gather data below the root"""
    rng = np.random.default_rng(11)
    t = Trace()
    a = t.input(F32, 32, 24)
    b = t.input(F32, 24, 32)
    c = t.input(F32, 24, 24)
    o1 = t.assign(t.output(F32, 32, 24), a - mean(a) + sum_(a * f32(0.5)))
    o2 = t.assign(t.output(F32, 24, 24), matmul(a * f32(2.0), b) + c)
    o3 = t.assign(t.output(F32, 24, 24), matmul((b + f32(1.0)).T, a.T) - f32(3.0))
    o4 = t.assign(t.output(F32, 32, 24), gather1_floor(a + f32(1.0), Trace.index(I32, 0, 32, 24) + 1))
    A = rng.integers(-4, 5, size=(1, 24, 32)).astype(np.float32)
    B = rng.integers(-4, 5, size=(1, 32, 24)).astype(np.float32)
    Cc = rng.integers(-4, 5, size=(1, 24, 24)).astype(np.float32)
    a2, b2 = A[0], B[0]
    e1 = (a2 - np.float32(a2.astype(np.float64).mean())) + np.float32((a2 * np.float32(0.5)).astype(np.float64).sum())
    e2 = (a2 * 2) @ b2 + Cc[0]
    e3 = (b2 + 1).T @ a2.T - 3
    e4 = np.roll(a2 + 1, -1, axis=1)
    return t, {a.index: A, b.index: B, c.index: Cc}, (o1, o2, o3, o4), (e1, e2, e3, e4)
