"""CPU oracle (TEST INFRASTRUCTURE ONLY).

ctypes face of liboracle.so (oracle.hpp: a restatement of the reference's CPU semantics plus brute-force checks) and,
when it was built, of _ref/libref.so (the reference's own CPU code compiled from its headers). Only tests/, the smoke
check and the CPU-baseline leg of bench.py use this module; the product package never imports it.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_PATH = os.path.join(_HERE, "liboracle.so")
REF_PATH = os.path.join(_HERE, "_ref", "libref.so")

_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_vp = ctypes.c_void_p

_ol = None
_rl = None


def _p(a, t):
    return a.ctypes.data_as(t)


def olib():
    global _ol
    if _ol is None:
        if not os.path.exists(ORACLE_PATH):
            raise RuntimeError(f"{ORACLE_PATH} missing: run `make -C oracle`")
        L = ctypes.CDLL(ORACLE_PATH)
        sig = {
            "orc_graph_from_pairs": (_vp, [_u32p, ctypes.c_uint64, ctypes.c_int64]),
            "orc_graph_load_text": (_vp, [ctypes.c_char_p]),
            "orc_graph_from_csr": (_vp, [ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p]),
            "orc_graph_free": (None, [_vp]),
            "orc_graph_n": (ctypes.c_uint32, [_vp]),
            "orc_graph_m": (ctypes.c_uint32, [_vp]),
            "orc_graph_max_out_degree": (ctypes.c_uint32, [_vp]),
            "orc_graph_copy": (None, [_vp, _u32p, _u32p, _u32p]),
            "orc_triangles_serial": (ctypes.c_uint64, [_vp]),
            "orc_triangles_omp": (ctypes.c_uint64, [_vp, ctypes.c_int]),
            "orc_omp_max_threads": (ctypes.c_int, []),
            "orc_participation": (ctypes.c_uint64, [_vp, _u8p, _u32p, ctypes.c_int]),
            "orc_edge_support": (ctypes.c_uint64, [_vp, _u32p]),
            "orc_batch_peel": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, _u64p, _u32p, ctypes.c_char_p, ctypes.c_int]),
            "orc_peel_arith_ok": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]),
            "orc_undirected": (None, [_vp, _u32p, _u32p]),
            "orc_core_decomposition": (ctypes.c_int, [_vp, _u32p, _u32p]),
            "orc_core_levels": (ctypes.c_int, [_vp, _u32p, _u32p, _u32p]),
            "orc_core_fixpoint": (None, [_vp, _u32p]),
            "orc_core_level_density": (None, [_vp, _u32p, ctypes.c_uint32, _u64p, _u64p]),
            "orc_edge_order": (None, [_vp, _u32p]),
            "orc_build_tasks": (ctypes.c_uint64, [_vp, ctypes.c_uint32, _u32p, _u32p, ctypes.c_uint64]),
            "orc_induced_subgraph": (_vp, [_vp, _u8p, _u32p, _u32p]),
            "orc_prune_to_core": (_vp, [_vp, _u32p, ctypes.c_uint64, _u32p, _u32p, _u32p]),
            "orc_truss": (ctypes.c_uint32, [_vp, _u32p]),
            "orc_greedy_pp": (None, [_vp, ctypes.c_uint32, _u64p, _u32p, _u64p]),
            "orc_brute_truss": (None, [ctypes.c_uint32, _u32p, ctypes.c_uint64, _u32p]),
            "orc_brute_participation": (ctypes.c_uint64, [ctypes.c_uint32, _u32p, ctypes.c_uint64, _u8p, _u32p]),
            "orc_brute_triangles": (ctypes.c_uint64, [ctypes.c_uint32, _u32p, ctypes.c_uint64]),
            "orc_brute_densest": (None, [ctypes.c_uint32, _u32p, ctypes.c_uint64, _u64p, _u64p]),
            "orc_brute_core": (None, [ctypes.c_uint32, _u32p, ctypes.c_uint64, _u32p]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _ol = L
    return _ol


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def rlib():
    global _rl
    if _rl is None:
        if not ref_available():
            raise RuntimeError(f"{REF_PATH} missing (the reference sources are needed to build it)")
        L = ctypes.CDLL(REF_PATH)
        sig = {
            "ref_load_graph": (_vp, [ctypes.c_char_p]),
            "ref_build_graph": (_vp, [ctypes.c_uint32, _u32p, ctypes.c_uint64]),
            "ref_free": (None, [_vp]),
            "ref_n": (ctypes.c_uint32, [_vp]),
            "ref_m": (ctypes.c_uint32, [_vp]),
            "ref_max_out_degree": (ctypes.c_uint32, [_vp]),
            "ref_copy": (None, [_vp, _u32p, _u32p, _u32p]),
            "ref_triangles_serial": (ctypes.c_uint64, [_vp]),
            "ref_triangles_omp": (ctypes.c_uint64, [_vp, ctypes.c_int]),
            "ref_participation": (ctypes.c_uint64, [_vp, _u8p, _u32p]),
            "ref_edge_support": (ctypes.c_uint64, [_vp, _u32p]),
            "ref_batch_peel": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, _u64p, _u32p]),
            "ref_core_decomposition": (ctypes.c_int, [_vp, _u32p, _u32p]),
            "ref_core_level_density": (None, [_vp, _u32p, ctypes.c_uint32, _u64p, _u64p]),
            "ref_edge_order": (None, [_vp, _u32p]),
            "ref_build_tasks": (ctypes.c_uint64, [_vp, ctypes.c_uint32, _u32p, _u32p, ctypes.c_uint64]),
            "ref_undirected": (None, [_vp, _u32p, _u32p]),
            "ref_peel_arith_ok": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]),
            "ref_induced_subgraph": (_vp, [_vp, _u8p, _u32p, _u32p]),
            "ref_prune_to_core": (_vp, [_vp, _u32p, ctypes.c_uint64, _u32p, _u32p, _u32p]),
            "ref_greedy_pp": (None, [_vp, ctypes.c_uint32, _u64p, _u32p, _u64p]),
            "ref_core_exact": (None, [_vp, _u32p, ctypes.c_int, _u64p, _u32p, _u64p, _u32p]),
            "ref_exact_densest": (None, [_vp, _u64p, _u32p]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _rl = L
    return _rl



def _greedy(fn, h, n, passes):
    out = np.zeros(4, np.uint64)
    st = np.zeros(max(n, 1), np.uint32)
    pp = np.zeros(2 * passes, np.uint64)
    fn(h, int(passes), _p(out, _u64p), _p(st, _u32p), _p(pp, _u64p))
    if not out[3]:
        raise RuntimeError("greedy_pp: triangle bookkeeping inconsistent")
    return int(out[0]), int(out[1]), st[: int(out[2])], [(int(a), int(b)) for a, b in pp.reshape(-1, 2)]


def _dsg_dict(out, st):
    return {"tri": int(out[0]), "verts": int(out[1]), "set": [int(x) for x in st[: int(out[2])]], "ok": bool(out[3]),
            "flow_calls": int(out[4]), "components_searched": int(out[5]), "shrinks": int(out[6])}


class OGraph:
    """Oracle graph (orc::Graph)."""

    def __init__(self, h):
        if not h:
            raise ValueError("oracle graph construction failed")
        self._h = h
        L = olib()
        self.n = L.orc_graph_n(h)
        self.m = L.orc_graph_m(h)
        self.max_out_degree = L.orc_graph_max_out_degree(h)

    @classmethod
    def from_pairs(cls, pairs: np.ndarray, n: Optional[int] = None) -> "OGraph":
        p = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1)
        return cls(olib().orc_graph_from_pairs(_p(p, _u32p), p.size // 2, -1 if n is None else n))

    @classmethod
    def load_text(cls, path: str) -> "OGraph":
        return cls(olib().orc_graph_load_text(path.encode()))

    @classmethod
    def from_csr(cls, rowptr, col) -> "OGraph":
        r = np.ascontiguousarray(rowptr, dtype=np.uint32)
        c = np.ascontiguousarray(col, dtype=np.uint32)
        return cls(olib().orc_graph_from_csr(r.size - 1, c.size, _p(r, _u32p), _p(c, _u32p)))

    def arrays(self):
        r = np.zeros(self.n + 1, np.uint32)
        c = np.zeros(max(self.m, 1), np.uint32)
        s = np.zeros(max(self.m, 1), np.uint32)
        olib().orc_graph_copy(self._h, _p(r, _u32p), _p(c, _u32p), _p(s, _u32p))
        return r, c[: self.m], s[: self.m]

    def triangles(self) -> int:
        return olib().orc_triangles_serial(self._h)

    def triangles_omp(self, threads: int = 0) -> int:
        return olib().orc_triangles_omp(self._h, threads)

    def participation(self, alive=None, omp=False):
        a = np.ones(self.n, np.uint8) if alive is None else np.ascontiguousarray(alive, dtype=np.uint8)
        tri = np.zeros(max(self.n, 1), np.uint32)
        T = olib().orc_participation(self._h, _p(a, _u8p), _p(tri, _u32p), 1 if omp else 0)
        return T, tri[: self.n]

    def edge_support(self):
        s = np.zeros(max(self.m, 1), np.uint32)
        T = olib().orc_edge_support(self._h, _p(s, _u32p))
        return T, s[: self.m]

    def batch_peel(self, num=1, den=10, omp=False):
        """Returns (best_num, best_den, best_round, rounds, removed_round) or raises ValueError(reason)."""
        st = np.zeros(4, np.uint64)
        rr = np.zeros(max(self.n, 1), np.uint32)
        err = ctypes.create_string_buffer(256)
        ok = olib().orc_batch_peel(self._h, num, den, 1 if omp else 0, _p(st, _u64p), _p(rr, _u32p), err, 256)
        if not ok:
            raise ValueError(err.value.decode())
        return int(st[0]), int(st[1]), int(st[2]), int(st[3]), rr[: self.n]

    def undirected(self):
        r = np.zeros(self.n + 1, np.uint32)
        c = np.zeros(max(2 * self.m, 1), np.uint32)
        olib().orc_undirected(self._h, _p(r, _u32p), _p(c, _u32p))
        return r, c[: 2 * self.m]

    def truss(self):
        """k-truss number of every oriented edge (index into col) and kmax (Wang & Cheng peeling)."""
        t = np.zeros(max(self.m, 1), np.uint32)
        kmax = olib().orc_truss(self._h, _p(t, _u32p))
        return t[: self.m], int(kmax)

    def greedy_pp(self, passes=1):
        """Greedy++ (dsg.h:189-250): (tri, verts, set ascending, per_pass [(tri, verts)] after each pass)."""
        return _greedy(olib().orc_greedy_pp, self._h, self.n, passes)

    def induced(self, keep):
        """induced_subgraph (dsg.h:103-114): (OGraph, old_id)."""
        k = np.ascontiguousarray(keep, dtype=np.uint8)
        old = np.zeros(max(self.n, 1), np.uint32)
        n2 = ctypes.c_uint32()
        h = olib().orc_induced_subgraph(self._h, _p(k, _u8p), _p(old, _u32p), ctypes.byref(n2))
        return OGraph(h), old[: n2.value]

    def prune_to_core(self, core, k):
        """prune_to_core (dsg.h:636-645): (OGraph, old_id, core_out)."""
        c = np.ascontiguousarray(core, dtype=np.uint32)
        old = np.zeros(max(self.n, 1), np.uint32)
        cout = np.zeros(max(self.n, 1), np.uint32)
        n2 = ctypes.c_uint32()
        h = olib().orc_prune_to_core(self._h, _p(c, _u32p), int(k), _p(old, _u32p), _p(cout, _u32p), ctypes.byref(n2))
        return OGraph(h), old[: n2.value], cout[: n2.value]

    def core(self):
        c = np.zeros(max(self.n, 1), np.uint32)
        k = ctypes.c_uint32()
        if not olib().orc_core_decomposition(self._h, _p(c, _u32p), ctypes.byref(k)):
            raise ValueError("core number does not fit uint32")
        return c[: self.n], k.value

    def core_levels(self):
        c = np.zeros(max(self.n, 1), np.uint32)
        r, k = ctypes.c_uint32(), ctypes.c_uint32()
        olib().orc_core_levels(self._h, _p(c, _u32p), ctypes.byref(r), ctypes.byref(k))
        return c[: self.n], r.value, k.value

    def core_fixpoint(self):
        c = np.zeros(max(self.n, 1), np.uint32)
        olib().orc_core_fixpoint(self._h, _p(c, _u32p))
        return c[: self.n]

    def core_level_density(self, core, k):
        c = np.ascontiguousarray(core, dtype=np.uint32)
        t, v = ctypes.c_uint64(), ctypes.c_uint64()
        olib().orc_core_level_density(self._h, _p(c, _u32p), k, ctypes.byref(t), ctypes.byref(v))
        return t.value, v.value

    def edge_order(self):
        p = np.zeros(max(self.m, 1), np.uint32)
        olib().orc_edge_order(self._h, _p(p, _u32p))
        return p[: self.m]

    def build_tasks(self, chunk=128):
        k = int(olib().orc_build_tasks(self._h, chunk, None, None, 0))
        te, ts = np.zeros(max(k, 1), np.uint32), np.zeros(max(k, 1), np.uint32)
        olib().orc_build_tasks(self._h, chunk, _p(te, _u32p), _p(ts, _u32p), k)
        return te[:k], ts[:k]

    def __del__(self):
        try:
            if self._h:
                olib().orc_graph_free(self._h)
                self._h = None
        except Exception:
            pass


class RGraph:
    """The reference's own Graph, built by the reference's own code (oracle/_ref)."""

    def __init__(self, h):
        if not h:
            raise ValueError("reference graph construction failed")
        self._h = h
        L = rlib()
        self.n, self.m, self.max_out_degree = L.ref_n(h), L.ref_m(h), L.ref_max_out_degree(h)

    @classmethod
    def load_text(cls, path: str) -> "RGraph":
        return cls(rlib().ref_load_graph(path.encode()))

    @classmethod
    def build(cls, n: int, unique_pairs: np.ndarray) -> "RGraph":
        p = np.ascontiguousarray(unique_pairs, dtype=np.uint32).reshape(-1)
        return cls(rlib().ref_build_graph(n, _p(p, _u32p), p.size // 2))

    def arrays(self):
        r = np.zeros(self.n + 1, np.uint32)
        c = np.zeros(max(self.m, 1), np.uint32)
        s = np.zeros(max(self.m, 1), np.uint32)
        rlib().ref_copy(self._h, _p(r, _u32p), _p(c, _u32p), _p(s, _u32p))
        return r, c[: self.m], s[: self.m]

    def triangles(self) -> int:
        return rlib().ref_triangles_serial(self._h)

    def triangles_omp(self, threads: int = 0) -> int:
        return rlib().ref_triangles_omp(self._h, threads)

    def participation(self, alive=None):
        a = np.ones(self.n, np.uint8) if alive is None else np.ascontiguousarray(alive, dtype=np.uint8)
        tri = np.zeros(max(self.n, 1), np.uint32)
        T = rlib().ref_participation(self._h, _p(a, _u8p), _p(tri, _u32p))
        return T, tri[: self.n]

    def edge_support(self):
        s = np.zeros(max(self.m, 1), np.uint32)
        T = rlib().ref_edge_support(self._h, _p(s, _u32p))
        return T, s[: self.m]

    def batch_peel(self, num=1, den=10, omp=False):
        st = np.zeros(4, np.uint64)
        rr = np.zeros(max(self.n, 1), np.uint32)
        if not rlib().ref_batch_peel(self._h, num, den, 1 if omp else 0, _p(st, _u64p), _p(rr, _u32p)):
            raise ValueError("reference batch peel failed")
        return int(st[0]), int(st[1]), int(st[2]), int(st[3]), rr[: self.n]

    def induced(self, keep):
        k = np.ascontiguousarray(keep, dtype=np.uint8)
        old = np.zeros(max(self.n, 1), np.uint32)
        n2 = ctypes.c_uint32()
        h = rlib().ref_induced_subgraph(self._h, _p(k, _u8p), _p(old, _u32p), ctypes.byref(n2))
        return RGraph(h), old[: n2.value]

    def greedy_pp(self, passes=1):
        return _greedy(rlib().ref_greedy_pp, self._h, self.n, passes)

    def core_exact(self, core, pruning12=True, seed=None):
        """Reference core_exact: dict of tri, verts, set, ok, flow_calls, components_searched, shrinks."""
        c = np.ascontiguousarray(core, dtype=np.uint32)
        out = np.zeros(7, np.uint64)
        st = np.zeros(max(self.n, 1), np.uint32)
        if seed is None:
            rlib().ref_core_exact(self._h, _p(c, _u32p), 1 if pruning12 else 0, None, None, _p(out, _u64p), _p(st, _u32p))
        else:
            tv = np.array([seed[0], seed[1]], np.uint64)
            ss = np.ascontiguousarray(seed[2], dtype=np.uint32)
            rlib().ref_core_exact(self._h, _p(c, _u32p), 1 if pruning12 else 0, _p(tv, _u64p), _p(ss, _u32p), _p(out, _u64p),
                                  _p(st, _u32p))
        return _dsg_dict(out, st)

    def exact_densest(self):
        out = np.zeros(7, np.uint64)
        st = np.zeros(max(self.n, 1), np.uint32)
        rlib().ref_exact_densest(self._h, _p(out, _u64p), _p(st, _u32p))
        return _dsg_dict(out, st)

    def prune_to_core(self, core, k):
        c = np.ascontiguousarray(core, dtype=np.uint32)
        old = np.zeros(max(self.n, 1), np.uint32)
        cout = np.zeros(max(self.n, 1), np.uint32)
        n2 = ctypes.c_uint32()
        h = rlib().ref_prune_to_core(self._h, _p(c, _u32p), int(k), _p(old, _u32p), _p(cout, _u32p), ctypes.byref(n2))
        return RGraph(h), old[: n2.value], cout[: n2.value]

    def core(self):
        c = np.zeros(max(self.n, 1), np.uint32)
        k = ctypes.c_uint32()
        if not rlib().ref_core_decomposition(self._h, _p(c, _u32p), ctypes.byref(k)):
            raise ValueError("core number does not fit uint32")
        return c[: self.n], k.value

    def core_level_density(self, core, k):
        c = np.ascontiguousarray(core, dtype=np.uint32)
        t, v = ctypes.c_uint64(), ctypes.c_uint64()
        rlib().ref_core_level_density(self._h, _p(c, _u32p), k, ctypes.byref(t), ctypes.byref(v))
        return t.value, v.value

    def edge_order(self):
        p = np.zeros(max(self.m, 1), np.uint32)
        rlib().ref_edge_order(self._h, _p(p, _u32p))
        return p[: self.m]

    def build_tasks(self, chunk=128):
        k = int(rlib().ref_build_tasks(self._h, chunk, None, None, 0))
        te, ts = np.zeros(max(k, 1), np.uint32), np.zeros(max(k, 1), np.uint32)
        rlib().ref_build_tasks(self._h, chunk, _p(te, _u32p), _p(ts, _u32p), k)
        return te[:k], ts[:k]

    def undirected(self):
        r = np.zeros(self.n + 1, np.uint32)
        c = np.zeros(max(2 * self.m, 1), np.uint32)
        rlib().ref_undirected(self._h, _p(r, _u32p), _p(c, _u32p))
        return r, c[: 2 * self.m]

    def __del__(self):
        try:
            if self._h:
                rlib().ref_free(self._h)
                self._h = None
        except Exception:
            pass


def canonical_unique(pairs: np.ndarray) -> np.ndarray:
    """(min, max) pairs, loops dropped, sorted unique: the edge set load_graph keeps (graph.h:85-92)."""
    p = np.asarray(pairs, dtype=np.uint32).reshape(-1, 2)
    a, b = p[:, 0], p[:, 1]
    keep = a != b
    lo = np.minimum(a, b)[keep].astype(np.uint64)
    hi = np.maximum(a, b)[keep].astype(np.uint64)
    key = np.unique((lo << np.uint64(32)) | hi)
    return np.stack([(key >> np.uint64(32)).astype(np.uint32), (key & np.uint64(0xFFFFFFFF)).astype(np.uint32)], axis=1)


def brute_triangles(n: int, unique_pairs: np.ndarray) -> int:
    p = np.ascontiguousarray(unique_pairs, dtype=np.uint32).reshape(-1)
    return olib().orc_brute_triangles(n, _p(p, _u32p), p.size // 2)


def brute_participation(n: int, unique_pairs: np.ndarray, alive: np.ndarray):
    p = np.ascontiguousarray(unique_pairs, dtype=np.uint32).reshape(-1)
    a = np.ascontiguousarray(alive, dtype=np.uint8)
    tri = np.zeros(max(n, 1), np.uint32)
    T = olib().orc_brute_participation(n, _p(p, _u32p), p.size // 2, _p(a, _u8p), _p(tri, _u32p))
    return T, tri[:n]


def brute_densest(n: int, unique_pairs: np.ndarray):
    p = np.ascontiguousarray(unique_pairs, dtype=np.uint32).reshape(-1)
    bn, bd = ctypes.c_uint64(), ctypes.c_uint64()
    olib().orc_brute_densest(n, _p(p, _u32p), p.size // 2, ctypes.byref(bn), ctypes.byref(bd))
    return bn.value, bd.value


def brute_core(n: int, unique_pairs: np.ndarray) -> np.ndarray:
    p = np.ascontiguousarray(unique_pairs, dtype=np.uint32).reshape(-1)
    c = np.zeros(max(n, 1), np.uint32)
    olib().orc_brute_core(n, _p(p, _u32p), p.size // 2, _p(c, _u32p))
    return c[:n]


def brute_truss(n: int, unique_pairs: np.ndarray) -> np.ndarray:
    """k-truss number of each undirected pair by the definition (small graphs)."""
    p = np.ascontiguousarray(unique_pairs, dtype=np.uint32).reshape(-1)
    t = np.zeros(max(p.size // 2, 1), np.uint32)
    olib().orc_brute_truss(n, _p(p, _u32p), p.size // 2, _p(t, _u32p))
    return t[: p.size // 2]


def peel_arith_ok(n, T0, num, den) -> bool:
    return bool(olib().orc_peel_arith_ok(n, T0, num, den))
