"""Pins the oracle's RNG block functions and stream layout against the
reference's own generator: the Random123 OpenCL text that Chai's JIT emits
(Random123/Random123.cpp:109-135, R123Philox.cpp:288-520,
R123Threefry.cpp:180-520), produced by the reference's emitter classes built
from their sources and compiled for the host as OpenCL C (`make -C oracle
ref` -> oracle/_ref/libchai_r123.so).

This is what pins Chai's own Threefry4x32-20, which mixes the word pairs
(0,1), (2,3) in every round (R123Threefry.cpp:238-255) instead of Random123's
alternating pairing, so no published vector covers it.  The mathematics adds
one more check: without the permutation the halves never mix, so output words
0 and 1 depend on counter words 0 and 1 (and the key) only, and words 2 and 3
on counter words 2 and 3 only."""
import ctypes
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def r123ref(oracle):
    L = oracle.r123_lib()
    if L is None:
        pytest.skip("oracle/_ref not built (the reference tree is not present here)")
    return L


def _call(fn, ctr, key, ctype, n):
    c = (ctype * len(ctr))(*ctr)
    k = (ctype * len(key))(*key)
    o = (ctype * n)()
    fn(c, k, o)
    return list(o)


def _inputs(words, kwords, bits, n, seed):
    rng = np.random.default_rng(seed)
    hi = 2 ** bits
    out = []
    for i in range(n):
        ctr = [int(x) for x in rng.integers(0, hi, size=words, dtype=np.uint64)]
        key = [int(x) for x in rng.integers(0, hi, size=kwords, dtype=np.uint64)]
        if i % 5 == 0:      # small counters and zero keys, as the stream uses
            ctr = [i] + [0] * (words - 1)
        if i % 7 == 0:
            key = [0] * kwords
        out.append((ctr, key))
    with open(os.path.join(GOLD, "r123_kat.json")) as f:
        return out, json.load(f)


CASES = [
    # oracle function, reference entry, counter words, key words, bits, KAT table
    ("threefry4x32", "chai_ref_threefry4x32", 4, 4, 32, "threefry4x32_20"),
    ("threefry2x32", "chai_ref_threefry2x32", 2, 2, 32, "threefry2x32_20"),
    ("threefry2x64", "chai_ref_threefry2x64", 2, 2, 64, "threefry2x64_20"),
    ("philox4x32", "chai_ref_philox4x32", 4, 2, 32, "philox4x32_10"),
    ("philox2x32", "chai_ref_philox2x32", 2, 1, 32, "philox2x32_10"),
]


@pytest.mark.parametrize("name,refname,words,kwords,bits,kat", CASES)
def test_block_function_matches_reference_text(oracle, r123ref, name, refname, words, kwords, bits, kat):
    ctype = ctypes.c_uint64 if bits == 64 else ctypes.c_uint32
    fn = getattr(r123ref, refname)
    mine = getattr(oracle, name)
    cases, katv = _inputs(words, kwords, bits, 3000, hash(name) % 1000)
    cases += [([int(x, 16) for x in e["ctr"]], [int(x, 16) for x in e["key"]]) for e in katv[kat]]
    for ctr, key in cases:
        assert mine(ctr, key) == _call(fn, ctr, key, ctype, words), (name, ctr, key)


def test_reference_threefry4x32_is_not_the_published_form(oracle, r123ref):
    """The emitted Threefry4x32-20 differs from Random123's KAT outputs (it is
    Chai's permutation-free variant) while Threefry2x32, which has one word
    pair, equals the published vectors."""
    with open(os.path.join(GOLD, "r123_kat.json")) as f:
        kat = json.load(f)
    for e in kat["threefry4x32_20"]:
        ctr, key = [int(x, 16) for x in e["ctr"]], [int(x, 16) for x in e["key"]]
        assert _call(r123ref.chai_ref_threefry4x32, ctr, key, ctypes.c_uint32, 4) != [int(x, 16) for x in e["out"]]
    for e in kat["threefry2x32_20"]:
        ctr, key = [int(x, 16) for x in e["ctr"]], [int(x, 16) for x in e["key"]]
        assert _call(r123ref.chai_ref_threefry2x32, ctr, key, ctypes.c_uint32, 2) == [int(x, 16) for x in e["out"]]


def test_threefry4x32_halves_are_independent(oracle):
    """Words 0,1 never see counter words 2,3 and vice versa: changing one
    half of the counter leaves the other half of the output unchanged, and
    changes its own half."""
    rng = np.random.default_rng(3)
    for _ in range(500):
        ctr = [int(x) for x in rng.integers(0, 2 ** 32, size=4, dtype=np.uint64)]
        key = [int(x) for x in rng.integers(0, 2 ** 32, size=4, dtype=np.uint64)]
        base = oracle.threefry4x32(ctr, key)
        hi = [int(x) for x in rng.integers(0, 2 ** 32, size=2, dtype=np.uint64)]
        lo = [int(x) for x in rng.integers(0, 2 ** 32, size=2, dtype=np.uint64)]
        a = oracle.threefry4x32(ctr[:2] + hi, key)
        b = oracle.threefry4x32(lo + ctr[2:], key)
        assert a[:2] == base[:2]
        assert b[2:] == base[2:]
        if hi != ctr[2:]:
            assert a[2:] != base[2:]
        if lo != ctr[:2]:
            assert b[:2] != base[:2]


@pytest.mark.parametrize("variant", [1230, 1231])
@pytest.mark.parametrize("vl", [1, 2, 4])
def test_stream_layout_matches_reference_wrappers(oracle, r123ref, variant, vl):
    """Key {vector index, 0xdecafbad, ...} and counter {seed, 0xf00dcafe, ...}
    as the reference's *_uniform_0_<vl>(tid, idx) wrappers build them
    (R123Philox.cpp:366-381, R123Threefry.cpp:377-387), including the
    one-word Philox2x32 key whose second initializer OpenCL C drops."""
    rng = np.random.default_rng(variant + vl)
    o = (ctypes.c_uint32 * 4)()
    pts = [(0, 0), (1, 7), (2 ** 32 - 1, 42)] + [tuple(int(x) for x in rng.integers(0, 2 ** 32, size=2))
                                                for _ in range(500)]
    for vidx, seed in pts:
        assert r123ref.chai_ref_stream_words(variant, vl, vidx, seed, o) == 0
        assert oracle.rng_words(variant, 32, vl, vidx, seed) == list(o)[:vl], (variant, vl, vidx, seed)


def test_threefry4x32_stream_words_differ_from_a_planted_mistake(oracle, r123ref):
    """A stream built with Random123's permuted rounds instead (a plausible
    restatement mistake) is caught by the reference text."""
    L = oracle.lib()
    o = (ctypes.c_uint32 * 4)()
    c = (ctypes.c_uint32 * 4)(7, 0xf00dcafe, 0xdeadbeef, 0xbeeff00d)
    k = (ctypes.c_uint32 * 4)(5, 0xdecafbad, 0xfacebead, 0x12345678)
    L.chai_or_threefry4x32_20_r123(c, k, o)
    assert r123ref.chai_ref_stream_words(1231, 4, 5, 7, (ctypes.c_uint32 * 4)()) == 0
    ref = (ctypes.c_uint32 * 4)()
    r123ref.chai_ref_stream_words(1231, 4, 5, 7, ref)
    assert list(o) != list(ref)
    assert oracle.rng_words(1231, 32, 4, 5, 7) == list(ref)
