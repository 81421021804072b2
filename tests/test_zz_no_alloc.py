"""ga_tc_count_noalloc (the count that allocates nothing; its partial sums live in the library's static device
scratch) called through the C ABI, checked bit for bit against ga_tc_count with the same variant and against the CPU
oracle: two calls in a row per variant, the second on a smaller graph, then more grid shapes, an empty graph and the
argument checks. The caller's output word is device memory allocated here with hipMalloc (the caller may allocate;
the entry point does not)."""
import ctypes

import numpy as np
import pytest

import graph_analyzer_amd as ga
import oracle
from graph_analyzer_amd import gen

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return _hip


def _noalloc():
    f = ga.lib().ga_tc_count_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_void_p]
    return f


class DeviceWord:
    def __init__(self):
        self.p = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(self.p), 8) == 0

    def poison(self):
        v = ctypes.c_uint64(0xDEADBEEFDEADBEEF)
        assert hip().hipMemcpy(self.p, ctypes.byref(v), 8, 1) == 0   # hipMemcpyHostToDevice

    def read(self):
        v = ctypes.c_uint64()
        assert hip().hipMemcpy(ctypes.byref(v), self.p, 8, 2) == 0   # hipMemcpyDeviceToHost
        return v.value

    def close(self):
        hip().hipFree(self.p)


def count_noalloc(ctx, g, variant, out, blocks=0):
    out.poison()
    rc = _noalloc()(ctx.handle, g.handle, variant, blocks, out.p)
    ctx.check(rc, "ga_tc_count_noalloc")
    ctx.sync()
    return out.read()


def test_noalloc_equals_count_two_calls_second_smaller(ctx):
    big = gen.chung_lu(60000, 600000, 2.1, 4.0, seed=41)
    small = gen.chung_lu(3000, 20000, 2.3, 3.0, seed=42)
    gb = ga.DeviceGraph.from_pairs(ctx, big)
    gs = ga.DeviceGraph.from_pairs(ctx, small)
    Tb, Ts = oracle.OGraph.from_pairs(big).triangles(), oracle.OGraph.from_pairs(small).triangles()
    out = DeviceWord()
    try:
        for v in (ga.TC_NOSRC, ga.TC_VERTEX):
            ref_b, ref_s = gb.count(v)[0], gs.count(v)[0]
            assert (ref_b, ref_s) == (Tb, Ts), v
            assert count_noalloc(ctx, gb, v, out) == ref_b, v
            assert count_noalloc(ctx, gs, v, out) == ref_s, v   # second call, smaller input, same scratch
            for blocks in (1, 3, 7, 256, 4096):
                assert count_noalloc(ctx, gs, v, out, blocks) == ref_s, (v, blocks)
                assert count_noalloc(ctx, gb, v, out, blocks) == ref_b, (v, blocks)
        empty = ga.DeviceGraph.from_pairs(ctx, np.zeros((0, 2), np.uint32))
        for v in (ga.TC_NOSRC, ga.TC_VERTEX):
            assert count_noalloc(ctx, empty, v, out) == 0 == empty.count(v)[0]
        empty.close()
    finally:
        out.close()
        gb.close()
        gs.close()


def test_noalloc_rejects_what_it_cannot_do(ctx):
    g = ga.DeviceGraph.from_pairs(ctx, gen.chung_lu(2000, 8000, 2.2, 3.0, seed=43))
    out = DeviceWord()
    try:
        f = _noalloc()
        assert f(ctx.handle, g.handle, ga.TC_NOSRC, 4097, out.p) == ga.GA_ERR_ARG      # beyond the static scratch
        assert b"4096" in ga.lib().ga_ctx_last_error(ctx.handle)
        assert f(ctx.handle, g.handle, ga.TC_HASH, 0, out.p) == ga.GA_ERR_ARG          # needs a count plan
        assert f(ctx.handle, g.handle, ga.TC_NOSRC, 0, None) == ga.GA_ERR_ARG          # no output word
        assert count_noalloc(ctx, g, ga.TC_NOSRC, out) == g.count(ga.TC_MERGE)[0]       # still fine afterwards
    finally:
        out.close()
        g.close()
