"""Seeded synthetic graphs for the five workloads (libga_gen.so, host C++ with OpenMP).

Every generator returns a raw (k, 2) uint32 pair array: duplicates, both orientations and self loops may occur; the
CSR builder canonicalises them exactly like the reference's text loader (graph.h:71-96). Outputs do not depend on
the OpenMP thread count.
"""
from __future__ import annotations

import ctypes
import math
import os

from typing import Optional

import numpy as np

from . import GEN_PATH


class _Edges(ctypes.Structure):
    _fields_ = [("pairs", ctypes.POINTER(ctypes.c_uint32)), ("count", ctypes.c_uint64), ("n", ctypes.c_uint32)]


_gl = None


def _gen():
    global _gl
    if _gl is None:
        if not os.path.exists(GEN_PATH):
            raise RuntimeError(f"{GEN_PATH} is missing: build it with __graft_entry__.build()")
        L = ctypes.CDLL(GEN_PATH)
        E = ctypes.POINTER(_Edges)
        L.ga_gen_chung_lu.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double, ctypes.c_double, ctypes.c_uint64, ctypes.c_int, E]
        L.ga_gen_chung_lu_exact.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double, ctypes.c_double, ctypes.c_uint64, E]
        L.ga_gen_chung_lu_offset.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double, ctypes.c_double]
        L.ga_gen_chung_lu_offset.restype = ctypes.c_double
        L.ga_gen_rmat.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_uint64, E]
        L.ga_gen_er.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64, E]
        L.ga_gen_chung_lu_alias.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double, ctypes.c_double, ctypes.c_uint64, ctypes.c_int, E]
        L.ga_gen_communities.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double, ctypes.c_uint64, E]
        L.ga_gen_write_text.argtypes = [E, ctypes.c_char_p, ctypes.c_char_p]
        L.ga_gen_free.argtypes = [E]
        L.ga_gen_free.restype = None
        P = ctypes.POINTER
        L.ga_gen_batch.argtypes = [ctypes.c_uint32, P(ctypes.c_uint32), P(ctypes.c_uint64), P(ctypes.c_int32), P(ctypes.c_double),
                                   P(ctypes.c_double), P(ctypes.c_uint64), P(ctypes.c_uint64), E]
        _gl = L
    return _gl


def _take(fn, *args) -> np.ndarray:
    e = _Edges()
    rc = fn(*args, ctypes.byref(e))
    if rc != 0:
        raise ValueError(f"generator failed with {rc}")
    try:
        if e.count == 0:
            return np.zeros((0, 2), np.uint32)
        a = np.ctypeslib.as_array(e.pairs, shape=(int(e.count) * 2,)).copy()
        return a.reshape(-1, 2)
    finally:
        _gen().ga_gen_free(ctypes.byref(e))


def write_text(pairs: np.ndarray, path: str, comment: Optional[str] = None):
    """SNAP-style text file, one "a<TAB>b" line per pair, after an optional '#' comment line."""
    p = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1)
    e = _Edges()
    e.pairs = p.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
    e.count = p.size // 2
    e.n = 0
    if _gen().ga_gen_write_text(ctypes.byref(e), os.fsencode(path), comment.encode() if comment else None) != 0:
        raise OSError(f"cannot write {path}")


def chung_lu(n: int, samples: int, gamma: float, i0: float = 0.0, seed: int = 1, permute: bool = True) -> np.ndarray:
    return _take(_gen().ga_gen_chung_lu, n, samples, gamma, i0, seed, 1 if permute else 0)


def chung_lu_exact(n: int, m: int, gamma: float, i0: float = 0.0, seed: int = 1) -> np.ndarray:
    """Exactly m unique loop-free undirected edges."""
    return _take(_gen().ga_gen_chung_lu_exact, n, m, gamma, i0, seed)


def chung_lu_offset(n: int, m: int, gamma: float, dmax: float) -> float:
    return _gen().ga_gen_chung_lu_offset(n, m, gamma, dmax)


def rmat(scale: int, samples: int, a=0.57, b=0.19, c=0.19, seed: int = 1) -> np.ndarray:
    return _take(_gen().ga_gen_rmat, scale, samples, a, b, c, seed)


def erdos_renyi(n: int, samples: int, seed: int = 1) -> np.ndarray:
    return _take(_gen().ga_gen_er, n, samples, seed)


def communities(n: int, count: int, smin: int, smax: int, p: float, seed: int = 1) -> np.ndarray:
    return _take(_gen().ga_gen_communities, n, count, smin, smax, p, seed)


# ---- the named workloads ----

def powerlaw_small(seed: int = 2101) -> np.ndarray:
    """Chung-Lu n=8,298, exactly m=100,762 unique edges, exponent 2.1, weight offset for max expected degree ~1,000."""
    n, m = 8298, 100762
    i0 = chung_lu_offset(n, m, 2.1, 1000.0)
    return chung_lu_exact(n, m, 2.1, i0, seed)


def rmat21(seed: int = 2102) -> np.ndarray:
    """R-MAT scale 21, (0.57, 0.19, 0.19, 0.05), edge factor 8: 16,777,216 raw samples."""
    return rmat(21, 8 << 21, 0.57, 0.19, 0.19, seed)


def powerlaw_communities(seed: int = 2103, n: int = 4_000_000, background: int = 15_200_000, count: int = 20_000) -> np.ndarray:
    """Chung-Lu n=4.0M, exponent 2.3 (no weight offset: the heaviest vertex ends near 160k neighbours) plus `count`
    planted communities of 10-200 vertices (log-uniform sizes) with intra-community edge probability 0.3. The
    communities alone give about 19.6M edges, so the background sample count is chosen so the de-duplicated total
    lands near 34.7M edges (seed 2103: n 3,999,999, m 34,762,082, T 93,996,589)."""
    bg = chung_lu(n, background, 2.3, 0.0, seed)
    cm = communities(n, count, 10, 200, 0.3, seed + 1)
    return np.concatenate([bg, cm])


def batch_mix_specs(count: int = 4096, seed: int = 2104, nmin: int = 1000, nmax: int = 200_000):
    """Per-graph parameters of the mix: n log-uniform in [nmin, nmax], mean degree uniform in [4, 64] (n * deg / 2
    sampled pairs), half Erdos-Renyi and half Chung-Lu (alias-table sampling) with exponent uniform in [2.0, 2.5]."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        deg = float(rng.uniform(4, 64))
        kind = "er" if i % 2 == 0 else "cl"
        gamma = float(rng.uniform(2.0, 2.5))
        specs.append({"n": n, "deg": deg, "kind": kind, "gamma": gamma, "seed": seed * 100003 + i})
    return specs


def _cl_dmax(spec) -> float:
    # cap the heaviest expected degree at 4 sqrt(n * deg) to keep counts within uint32 tri[] counters
    n, deg = spec["n"], spec["deg"]
    return min(n / 2.0, 4.0 * math.sqrt(n * deg))


def _cl_offset(spec) -> float:
    n, deg = spec["n"], spec["deg"]
    return chung_lu_offset(n, int(n * deg / 2), spec["gamma"], _cl_dmax(spec))


def batch_graph(spec) -> np.ndarray:
    n, deg = spec["n"], spec["deg"]
    samples = int(n * deg / 2)
    if spec["kind"] == "er":
        return erdos_renyi(n, samples, spec["seed"])
    return _take(_gen().ga_gen_chung_lu_alias, n, samples, spec["gamma"], _cl_offset(spec), spec["seed"], 1)


class BatchPairs:
    """All graphs of a batch mix in one generator-owned buffer (no copy): pairs is a (k, 2) uint32 view, offsets the
    count + 1 pair offsets. Keep this object alive while pairs is in use."""

    def __init__(self, specs):
        c = len(specs)
        n = np.array([s["n"] for s in specs], np.uint32)
        samples = np.array([int(s["n"] * s["deg"] / 2) for s in specs], np.uint64)
        kind = np.array([0 if s["kind"] == "er" else 2 for s in specs], np.int32)
        gamma = np.array([s["gamma"] for s in specs], np.float64)
        i0 = np.array([0.0 if s["kind"] == "er" else -_cl_dmax(s) for s in specs], np.float64)   # solved in C
        seed = np.array([s["seed"] for s in specs], np.uint64)
        self.offsets = np.zeros(c + 1, np.uint64)
        self._e = _Edges()
        P = ctypes.POINTER
        rc = _gen().ga_gen_batch(c, n.ctypes.data_as(P(ctypes.c_uint32)), samples.ctypes.data_as(P(ctypes.c_uint64)),
                                 kind.ctypes.data_as(P(ctypes.c_int32)), gamma.ctypes.data_as(P(ctypes.c_double)),
                                 i0.ctypes.data_as(P(ctypes.c_double)), seed.ctypes.data_as(P(ctypes.c_uint64)),
                                 self.offsets.ctypes.data_as(P(ctypes.c_uint64)), ctypes.byref(self._e))
        if rc != 0:
            raise ValueError(f"batch generator failed with {rc}")
        k = int(self._e.count)
        self.pairs = np.ctypeslib.as_array(self._e.pairs, shape=(max(k, 1) * 2,))[: 2 * k].reshape(-1, 2)

    def graph(self, i) -> np.ndarray:
        return self.pairs[int(self.offsets[i]): int(self.offsets[i + 1])]

    def __del__(self):
        try:
            self.pairs = None
            _gen().ga_gen_free(ctypes.byref(self._e))
        except Exception:
            pass
