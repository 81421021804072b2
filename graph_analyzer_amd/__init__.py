"""graph_analyzer_amd - MI355X-native triangle counting, triangle-density batch peeling and (k,triangle)-core.

Python face of the C ABI in include/ga_amd.h (libga_amd.so, HIP kernels for gfx950). The C++ face, which restates
the reference's graph.h / densest.h / dsg.h API, is in graph_analyzer_amd/host/. There is no CPU fallback: if the
HIP library is missing or no GPU is visible, every call raises.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libga_amd.so")
GEN_PATH = os.path.join(_HERE, "libga_gen.so")

GA_OK, GA_ERR_ARG, GA_ERR_HIP, GA_ERR_OOM, GA_ERR_ARITH, GA_ERR_STATE = range(6)

# triangle-count variants (ga_amd.h)
TC_AUTO, TC_MERGE, TC_GROUP4, TC_GROUP8, TC_GROUP16, TC_GROUP64, TC_HASH, TC_SORTED, TC_NOSRC, TC_VERTEX = range(10)
TC_VARIANTS = {
    "merge": TC_MERGE,
    "group4": TC_GROUP4,
    "group8": TC_GROUP8,
    "group16": TC_GROUP16,
    "group64": TC_GROUP64,
    "hash": TC_HASH,
    "sorted": TC_SORTED,
    "nosrc": TC_NOSRC,
    "vertex": TC_VERTEX,
}

_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_vp = ctypes.c_void_p


class GAError(RuntimeError):
    """A non-zero status from the C ABI; .code is the GA_ERR_* value."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _exact_dict(out, st):
    return {"tri": int(out[0]), "verts": int(out[1]), "set": [int(x) for x in st[: int(out[2])]], "ok": bool(out[3]),
            "flow_calls": int(out[4]), "components_searched": int(out[5]), "shrinks": int(out[6])}


class TcTuning(ctypes.Structure):
    _fields_ = [("lanes", ctypes.c_int32 * 5), ("block", ctypes.c_int32), ("work_chunk", ctypes.c_int32),
                ("batches", ctypes.c_int32), ("load_shift", ctypes.c_int32), ("hash_cap", ctypes.c_int32),
                ("flat_chunk", ctypes.c_int32), ("flat_block", ctypes.c_int32),
                ("item_order", ctypes.c_int32), ("relabel", ctypes.c_int32), ("probe", ctypes.c_int32),
                ("waves", ctypes.c_int32), ("streams", ctypes.c_int32)]

    @classmethod
    def from_dict(cls, d):
        t = cls()
        for k, v in d.items():
            if k == "lanes":
                t.lanes[:] = v
            else:
                setattr(t, k, v)
        return t

    def as_dict(self):
        return {"lanes": list(self.lanes), "block": self.block, "work_chunk": self.work_chunk, "batches": self.batches,
                "load_shift": self.load_shift, "hash_cap": self.hash_cap, "flat_chunk": self.flat_chunk,
                "flat_block": self.flat_block,
                "item_order": self.item_order, "relabel": self.relabel, "probe": self.probe,
                "waves": self.waves, "streams": self.streams}


_lib = None


def lib() -> ctypes.CDLL:
    """The loaded HIP library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() (there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "ga_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
            "ga_ctx_destroy": (None, [_vp]),
            "ga_ctx_last_error": (ctypes.c_char_p, [_vp]),
            "ga_ctx_sync": (ctypes.c_int, [_vp]),
            "ga_ctx_device_info": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                                  ctypes.POINTER(ctypes.c_int), _u64p]),
            "ga_ctx_copy_probe": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_float)]),
            "ga_ctx_gather_probe": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.c_uint32, _u64p, ctypes.POINTER(ctypes.c_float)]),
            "ga_graph_from_pairs": (ctypes.c_int, [_vp, _u32p, ctypes.c_uint64, ctypes.c_int64, ctypes.POINTER(_vp),
                                                   ctypes.POINTER(ctypes.c_float)]),
            "ga_graph_from_csr": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p, ctypes.POINTER(_vp)]),
            "ga_graph_destroy": (None, [_vp]),
            "ga_graph_info": (ctypes.c_int, [_vp, _u32p, _u32p, _u32p]),
            "ga_graph_download": (ctypes.c_int, [_vp, _vp, _u32p, _u32p, _u32p]),
            "ga_graph_undirected": (ctypes.c_int, [_vp, _vp, _u32p, _u32p]),
            "ga_graph_edge_order": (ctypes.c_int, [_vp, _vp, _u32p]),
            "ga_graph_tasks": (ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _u32p, _u32p, ctypes.c_uint64, _u64p]),
            "ga_graph_from_text": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_uint64, ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_float)]),
            "ga_graph_load_text": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_float)]),
            "ga_graph_save": (ctypes.c_int, [_vp, _vp, ctypes.c_char_p]),
            "ga_graph_load": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_float)]),
            "ga_graph_induced": (ctypes.c_int, [_vp, _vp, _u8p, ctypes.POINTER(_vp), _u32p, ctypes.POINTER(ctypes.c_float)]),
            "ga_graph_prune_to_core": (ctypes.c_int, [_vp, _vp, _u32p, ctypes.c_uint64, ctypes.POINTER(_vp), _u32p, _u32p,
                                                      ctypes.POINTER(ctypes.c_float)]),
            "ga_truss_create": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(_vp)]),
            "ga_truss_destroy": (None, [_vp]),
            "ga_truss_run": (ctypes.c_int, [_vp, _u32p, ctypes.POINTER(ctypes.c_float)]),
            "ga_truss_download": (ctypes.c_int, [_vp, _u32p]),
            "ga_greedy_pp": (ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _u64p, _u64p, _u64p, _u32p, _u32p, ctypes.POINTER(ctypes.c_float)]),
            "ga_greedy_max_vertices": (ctypes.c_uint32, [_vp]),
            "ga_greedy_pp_many": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.c_uint32, ctypes.c_uint32, _u64p, _u8p,
                                                 ctypes.POINTER(ctypes.c_float)]),
            "ga_core_prune": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.POINTER(_vp), _u32p, _u32p, ctypes.POINTER(ctypes.c_float)]),
            "ga_tc_variant_name": (ctypes.c_char_p, [ctypes.c_int]),
            "ga_tc_count": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _u64p, ctypes.POINTER(ctypes.c_float)]),
            "ga_tc_count_noalloc": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_uint32, _vp]),
            "ga_batch_work": (ctypes.c_int, [_vp, _vp, _u64p, ctypes.POINTER(ctypes.c_float)]),
            "ga_ctx_barrier_probe": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_float),
                                                    ctypes.POINTER(ctypes.c_float)]),
            "ga_tc_count_many": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.c_uint32, ctypes.c_int, _u64p,
                                                ctypes.POINTER(ctypes.c_float)]),
            "ga_ctx_mem_info": (ctypes.c_int, [_vp, _u64p, _u64p]),
            "ga_ctx_set_grid_cap": (ctypes.c_int, [_vp, ctypes.c_uint32]),
            "ga_peel_run_batch": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.c_uint32, _u64p, ctypes.c_uint64, ctypes.c_uint64,
                                                 _u64p, ctypes.POINTER(ctypes.c_float)]),
            "ga_peel_set_ctx": (ctypes.c_int, [_vp, _vp]),
            "ga_core_run_many": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.c_uint32, _u64p, _u32p, ctypes.POINTER(ctypes.c_float)]),
            "ga_core_run_batch": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.c_uint32, _u64p, _u32p, ctypes.POINTER(ctypes.c_float)]),
            "ga_core_set_ctx": (ctypes.c_int, [_vp, _vp]),
            "ga_tc_prepare": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(ctypes.c_float)]),
            "ga_tc_get_tuning": (ctypes.c_int, [_vp, ctypes.POINTER(TcTuning)]),
            "ga_tc_set_tuning": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(TcTuning)]),
            "ga_tc_default_tuning": (ctypes.c_int, [ctypes.POINTER(TcTuning)]),
            "ga_tc_tune": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(TcTuning), ctypes.POINTER(ctypes.c_float)]),
            "ga_tc_plan_info": (ctypes.c_int, [_vp, _vp, _u32p, _u32p]),
            "ga_tc_bucket_profile": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _u64p, _u64p, ctypes.POINTER(ctypes.c_float)]),
            "ga_core_exact": (ctypes.c_int, [_vp, _vp, _u32p, ctypes.c_int, _u64p, _u32p, _u64p, _u32p]),
            "ga_exact_densest": (ctypes.c_int, [_vp, _vp, _u64p, _u32p]),
            "ga_tc_vertex": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _u8p, _u32p, _u64p]),
            "ga_tc_edge_support": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _u32p, _u64p]),
            "ga_peel_create": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(_vp)]),
            "ga_peel_destroy": (None, [_vp]),
            "ga_peel_reset_state": (ctypes.c_int, [_vp]),
            "ga_peel_begin_round": (ctypes.c_int, [_vp]),
            "ga_peel_run_tc": (ctypes.c_int, [_vp]),
            "ga_peel_run_mark": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32]),
            "ga_peel_read_stats": (ctypes.c_int, [_vp, _u64p, _u64p]),
            "ga_peel_download_removed_round": (ctypes.c_int, [_vp, _u32p]),
            "ga_peel_run": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, _u64p,
                                           ctypes.POINTER(ctypes.c_float)]),
            "ga_core_create": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.POINTER(_vp)]),
            "ga_core_destroy": (None, [_vp]),
            "ga_core_reset_state": (ctypes.c_int, [_vp]),
            "ga_core_begin_round": (ctypes.c_int, [_vp]),
            "ga_core_run_tc": (ctypes.c_int, [_vp]),
            "ga_core_run_min": (ctypes.c_int, [_vp]),
            "ga_core_run_peel": (ctypes.c_int, [_vp, ctypes.c_uint64]),
            "ga_core_read_stats": (ctypes.c_int, [_vp, _u64p, _u64p, _u64p]),
            "ga_core_init_tri": (ctypes.c_int, [_vp, _u64p]),
            "ga_core_run_find": (ctypes.c_int, [_vp, ctypes.c_uint64]),
            "ga_core_run_update": (ctypes.c_int, [_vp]),
            "ga_core_run_retire": (ctypes.c_int, [_vp]),
            "ga_core_download": (ctypes.c_int, [_vp, _u32p]),
            "ga_core_run": (ctypes.c_int, [_vp, ctypes.c_uint64, _u32p, ctypes.POINTER(ctypes.c_float)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


class Context:
    """One HIP device + stream (ga_ctx)."""

    def __init__(self, device: int = 0):
        self._h = _vp()
        rc = lib().ga_ctx_create(device, ctypes.byref(self._h))
        if rc != GA_OK:
            raise GAError(rc, f"ga_ctx_create({device}) failed (no HIP device?)")

    @property
    def handle(self):
        return self._h

    def check(self, rc: int, what: str = ""):
        if rc != GA_OK:
            msg = lib().ga_ctx_last_error(self._h)
            raise GAError(rc, f"{what}: {msg.decode() if msg else ''}")

    def sync(self):
        self.check(lib().ga_ctx_sync(self._h), "sync")

    def device_info(self) -> dict:
        name = ctypes.create_string_buffer(256)
        cus, wave, hbm = ctypes.c_int(), ctypes.c_int(), ctypes.c_uint64()
        self.check(lib().ga_ctx_device_info(self._h, name, 256, ctypes.byref(cus), ctypes.byref(wave), ctypes.byref(hbm)), "device_info")
        return {"name": name.value.decode(), "cus": cus.value, "wave": wave.value, "hbm_bytes": hbm.value}

    def greedy_max_vertices(self) -> int:
        """The largest graph (vertices) DeviceGraph.greedy_pp takes on this device: one workgroup's LDS holds its state."""
        return int(lib().ga_greedy_max_vertices(self._h))

    def barrier_probe(self, workgroups: int = 256, iters: int = 2000):
        """(us per device-wide barrier in a cooperative kernel, us per back-to-back launch) at this grid size."""
        b, l = ctypes.c_float(), ctypes.c_float()
        self.check(lib().ga_ctx_barrier_probe(self._h, workgroups, iters, ctypes.byref(b), ctypes.byref(l)), "barrier_probe")
        return b.value, l.value

    def set_grid_cap(self, max_workgroups: int = 0):
        """Caps every count, participation, support, peel and core launch of this context at max_workgroups workgroups
        (0: the library's own grids), so that tests can run the grid-stride and item loops many passes each."""
        self.check(lib().ga_ctx_set_grid_cap(self._h, int(max_workgroups)), "set_grid_cap")

    def mem_info(self):
        """(free, total) device bytes."""
        f, t = ctypes.c_uint64(), ctypes.c_uint64()
        self.check(lib().ga_ctx_mem_info(self._h, ctypes.byref(f), ctypes.byref(t)), "mem_info")
        return f.value, t.value

    def count_many(self, graphs, variant: int = 0):
        """Exact counts of several resident graphs on this context's stream with one readback: for the hash variant one
        launch per bucket shape covers every graph (graphs on the flat kernel, and other variants, get their own)."""
        hs = (_vp * max(len(graphs), 1))(*[g.handle for g in graphs])
        T = np.zeros(max(len(graphs), 1), np.uint64)
        ms = ctypes.c_float()
        self.check(lib().ga_tc_count_many(self._h, hs, len(graphs), variant, _ptr(T, _u64p), ctypes.byref(ms)), "tc_count_many")
        return T[: len(graphs)], ms.value

    def greedy_pp_many(self, graphs, passes: int = 1):
        """Greedy++ on several resident graphs, one workgroup per graph in one launch where they fit the LDS-bitmap form.
        Returns (stats, sets, ms): stats[i] = (tri, verts) of graph i's best density, sets[i] its vertices ascending."""
        hs = (_vp * max(len(graphs), 1))(*[g.handle for g in graphs])
        st = np.zeros(2 * max(len(graphs), 1), np.uint64)
        total = sum(g.n for g in graphs)
        flags = np.zeros(max(total, 1), np.uint8)
        ms = ctypes.c_float()
        self.check(lib().ga_greedy_pp_many(self._h, hs, len(graphs), int(passes), _ptr(st, _u64p), _ptr(flags, _u8p),
                                           ctypes.byref(ms)), "greedy_pp_many")
        sets, off = [], 0
        for g in graphs:
            sets.append(np.flatnonzero(flags[off: off + g.n]).astype(np.uint32))
            off += g.n
        return st[: 2 * len(graphs)].reshape(-1, 2), sets, ms.value

    def copy_probe(self, nbytes: int) -> float:
        """One plain device copy of nbytes (k_copy_probe); returns its device ms."""
        ms = ctypes.c_float()
        self.check(lib().ga_ctx_copy_probe(self._h, nbytes, ctypes.byref(ms)), "copy_probe")
        return ms.value

    def gather_probe(self, nbytes: int, seg_bytes: int):
        """Random reads of seg_bytes-long segments from an nbytes buffer (k_gather_probe); returns (segments, ms)."""
        segs = ctypes.c_uint64()
        ms = ctypes.c_float()
        self.check(lib().ga_ctx_gather_probe(self._h, nbytes, seg_bytes, ctypes.byref(segs), ctypes.byref(ms)), "gather_probe")
        return segs.value, ms.value

    def close(self):
        if self._h:
            lib().ga_ctx_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class PeelResult:
    """Mirror of PeelResult (densest.h:46-54)."""
    best_num: int
    best_den: int
    best_round: int
    rounds: int
    removed_round: np.ndarray

    def same(self, o: "PeelResult") -> bool:
        return (self.best_num == o.best_num and self.best_den == o.best_den and self.best_round == o.best_round
                and self.rounds == o.rounds and np.array_equal(self.removed_round, o.removed_round))


class DeviceGraph:
    """Oriented CSR resident in HBM (ga_graph)."""

    def __init__(self, ctx: Context, handle, build_ms: float = 0.0):
        self.ctx = ctx
        self._h = handle
        self.build_ms = build_ms
        n, m, mo = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        ctx.check(lib().ga_graph_info(self._h, ctypes.byref(n), ctypes.byref(m), ctypes.byref(mo)), "graph_info")
        self.n, self.m, self.max_out_degree = n.value, m.value, mo.value

    @classmethod
    def from_pairs(cls, ctx: Context, pairs: np.ndarray, n: Optional[int] = None) -> "DeviceGraph":
        """n=None: load_graph semantics on raw pairs; else build_graph_from_edges semantics (unique pairs, ids < n)."""
        p = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1)
        h = _vp()
        ms = ctypes.c_float()
        ctx.check(lib().ga_graph_from_pairs(ctx.handle, _ptr(p, _u32p), p.size // 2, -1 if n is None else int(n),
                                            ctypes.byref(h), ctypes.byref(ms)), "graph_from_pairs")
        return cls(ctx, h, ms.value)

    @classmethod
    def from_csr(cls, ctx: Context, rowptr: np.ndarray, col: np.ndarray) -> "DeviceGraph":
        r = np.ascontiguousarray(rowptr, dtype=np.uint32)
        c = np.ascontiguousarray(col, dtype=np.uint32)
        h = _vp()
        ctx.check(lib().ga_graph_from_csr(ctx.handle, r.size - 1, c.size, _ptr(r, _u32p), _ptr(c, _u32p), ctypes.byref(h)), "graph_from_csr")
        return cls(ctx, h)

    @classmethod
    def from_text(cls, ctx: Context, text: bytes) -> "DeviceGraph":
        """load_graph on an in-memory SNAP edge list, parsed on the GPU."""
        h = _vp()
        ms = ctypes.c_float()
        ctx.check(lib().ga_graph_from_text(ctx.handle, text, len(text), ctypes.byref(h), ctypes.byref(ms)), "graph_from_text")
        return cls(ctx, h, ms.value)

    @classmethod
    def load_text(cls, ctx: Context, path: str) -> "DeviceGraph":
        """load_graph (graph.h:71-96) with the text parse on the GPU."""
        h = _vp()
        ms = ctypes.c_float()
        ctx.check(lib().ga_graph_load_text(ctx.handle, os.fsencode(path), ctypes.byref(h), ctypes.byref(ms)), "graph_load_text")
        return cls(ctx, h, ms.value)

    @classmethod
    def load(cls, ctx: Context, path: str) -> "DeviceGraph":
        """A binary CSR file written by save(), checked on the device."""
        h = _vp()
        ms = ctypes.c_float()
        ctx.check(lib().ga_graph_load(ctx.handle, os.fsencode(path), ctypes.byref(h), ctypes.byref(ms)), "graph_load")
        return cls(ctx, h, ms.value)

    def save(self, path: str):
        """Writes the binary CSR file (GACSR001: 64-byte header, rowptr, col)."""
        self.ctx.check(lib().ga_graph_save(self.ctx.handle, self._h, os.fsencode(path)), "graph_save")

    @property
    def handle(self):
        return self._h

    def download(self):
        rowptr = np.zeros(self.n + 1, np.uint32)
        col = np.zeros(max(self.m, 1), np.uint32)
        src = np.zeros(max(self.m, 1), np.uint32)
        self.ctx.check(lib().ga_graph_download(self.ctx.handle, self._h, _ptr(rowptr, _u32p), _ptr(col, _u32p), _ptr(src, _u32p)), "download")
        return rowptr, col[: self.m], src[: self.m]

    def undirected(self):
        urow = np.zeros(self.n + 1, np.uint32)
        ucol = np.zeros(max(2 * self.m, 1), np.uint32)
        self.ctx.check(lib().ga_graph_undirected(self.ctx.handle, self._h, _ptr(urow, _u32p), _ptr(ucol, _u32p)), "undirected")
        return urow, ucol[: 2 * self.m]

    def edge_order(self) -> np.ndarray:
        perm = np.zeros(max(self.m, 1), np.uint32)
        self.ctx.check(lib().ga_graph_edge_order(self.ctx.handle, self._h, _ptr(perm, _u32p)), "edge_order")
        return perm[: self.m]

    def build_tasks(self, chunk: int = 128):
        """build_tasks (graph.h:138-150) on the device: (t_edge, t_start), every edge's min(d+u, d+v) in chunks."""
        nt = ctypes.c_uint64()
        self.ctx.check(lib().ga_graph_tasks(self.ctx.handle, self._h, chunk, None, None, 0, ctypes.byref(nt)), "graph_tasks")
        k = nt.value
        te, ts = np.zeros(max(k, 1), np.uint32), np.zeros(max(k, 1), np.uint32)
        if k:
            self.ctx.check(lib().ga_graph_tasks(self.ctx.handle, self._h, chunk, _ptr(te, _u32p), _ptr(ts, _u32p), k,
                                                ctypes.byref(nt)), "graph_tasks")
        return te[:k], ts[:k]

    # ---- subgraphs (dsg.h:103-114, 636-645) ----
    def induced(self, keep: np.ndarray):
        """induced_subgraph on {v : keep[v] != 0}: returns (DeviceGraph, old_id); new ids follow the old ids' order."""
        k = np.ascontiguousarray(keep, dtype=np.uint8)
        if k.size != self.n:
            raise ValueError(f"keep has {k.size} entries, the graph {self.n} vertices")
        old = np.zeros(max(self.n, 1), np.uint32)
        h = _vp()
        ms = ctypes.c_float()
        self.ctx.check(lib().ga_graph_induced(self.ctx.handle, self._h, _ptr(k, _u8p) if k.size else None, ctypes.byref(h),
                                              _ptr(old, _u32p), ctypes.byref(ms)), "graph_induced")
        sub = DeviceGraph(self.ctx, h, ms.value)
        return sub, old[: sub.n]

    def prune_to_core(self, core: np.ndarray, k: int):
        """prune_to_core: the subgraph on {v : core[v] >= k}; returns (DeviceGraph, old_id, core_out)."""
        c = np.ascontiguousarray(core, dtype=np.uint32)
        if c.size != self.n:
            raise ValueError(f"core has {c.size} entries, the graph {self.n} vertices")
        old = np.zeros(max(self.n, 1), np.uint32)
        cout = np.zeros(max(self.n, 1), np.uint32)
        h = _vp()
        ms = ctypes.c_float()
        self.ctx.check(lib().ga_graph_prune_to_core(self.ctx.handle, self._h, _ptr(c, _u32p) if c.size else None, int(k), ctypes.byref(h),
                                                    _ptr(old, _u32p), _ptr(cout, _u32p), ctypes.byref(ms)), "graph_prune_to_core")
        sub = DeviceGraph(self.ctx, h, ms.value)
        return sub, old[: sub.n], cout[: sub.n]

    # ---- triangle counting ----
    def count(self, variant: int = TC_AUTO):
        T = ctypes.c_uint64()
        ms = ctypes.c_float()
        self.ctx.check(lib().ga_tc_count(self.ctx.handle, self._h, variant, ctypes.byref(T), ctypes.byref(ms)), f"tc_count[{variant}]")
        return T.value, ms.value

    def prepare(self) -> float:
        ms = ctypes.c_float()
        self.ctx.check(lib().ga_tc_prepare(self.ctx.handle, self._h, ctypes.byref(ms)), "tc_prepare")
        return ms.value

    def plan_info(self):
        nw = ctypes.c_uint32()
        it = np.zeros(6, np.uint32)
        self.ctx.check(lib().ga_tc_plan_info(self.ctx.handle, self._h, ctypes.byref(nw), _ptr(it, _u32p)), "tc_plan_info")
        return nw.value, it.tolist()

    def core_exact(self, core, pruning12: bool = True, seed=None):
        """CoreExact (dsg.h:454-627) on the core numbers `core`; seed = (tri, verts, vertex set) or None. Returns a dict
        of tri, verts, set, flow_calls, components_searched, shrinks (GAError when a self-check fails)."""
        c = np.ascontiguousarray(core, dtype=np.uint32)
        if c.size != self.n:
            raise ValueError("core must have n entries")
        out = np.zeros(7, np.uint64)
        st = np.zeros(max(self.n, 1), np.uint32)
        tv = ss = None
        if seed is not None:
            tv = np.array([seed[0], seed[1]], np.uint64)
            ss = np.ascontiguousarray(seed[2], dtype=np.uint32)
        self.ctx.check(lib().ga_core_exact(self.ctx.handle, self._h, _ptr(c, _u32p), 1 if pruning12 else 0,
                                           None if tv is None else _ptr(tv, _u64p), None if ss is None else _ptr(ss, _u32p),
                                           _ptr(out, _u64p), _ptr(st, _u32p)), "core_exact")
        return _exact_dict(out, st)

    def exact_densest(self):
        """Exact (dsg.h:402-428): Alg 1 on the whole graph; the same dict as core_exact."""
        out = np.zeros(7, np.uint64)
        st = np.zeros(max(self.n, 1), np.uint32)
        self.ctx.check(lib().ga_exact_densest(self.ctx.handle, self._h, _ptr(out, _u64p), _ptr(st, _u32p)), "exact_densest")
        return _exact_dict(out, st)

    def bucket_profile(self, reps: int = 20):
        """Per launch of the plain count (5 hashed buckets, the binary-search bucket, the flat kernel): work entries,
        probe elements and mean device ms over reps counts."""
        ent = np.zeros(7, np.uint64)
        el = np.zeros(7, np.uint64)
        ms = (ctypes.c_float * 7)()
        self.ctx.check(lib().ga_tc_bucket_profile(self.ctx.handle, self._h, int(reps), _ptr(ent, _u64p), _ptr(el, _u64p), ms),
                       "tc_bucket_profile")
        return ent.tolist(), el.tolist(), list(ms)

    def get_tuning(self) -> TcTuning:
        t = TcTuning()
        self.ctx.check(lib().ga_tc_get_tuning(self._h, ctypes.byref(t)), "get_tuning")
        return t

    def set_tuning(self, t: TcTuning):
        self.ctx.check(lib().ga_tc_set_tuning(self.ctx.handle, self._h, ctypes.byref(t)), "set_tuning")

    def tune(self):
        t = TcTuning()
        ms = ctypes.c_float()
        self.ctx.check(lib().ga_tc_tune(self.ctx.handle, self._h, ctypes.byref(t), ctypes.byref(ms)), "tc_tune")
        return t, ms.value

    def vertex_participation(self, alive: Optional[np.ndarray] = None, variant: int = TC_AUTO):
        tri = np.zeros(max(self.n, 1), np.uint32)
        T = ctypes.c_uint64()
        ap = None
        if alive is not None:
            a = np.ascontiguousarray(alive, dtype=np.uint8)
            assert a.size == self.n
            ap = _ptr(a, _u8p)
        self.ctx.check(lib().ga_tc_vertex(self.ctx.handle, self._h, variant, ap, _ptr(tri, _u32p), ctypes.byref(T)), "tc_vertex")
        return T.value, tri[: self.n]

    def edge_support(self, variant: int = TC_AUTO):
        sup = np.zeros(max(self.m, 1), np.uint32)
        T = ctypes.c_uint64()
        self.ctx.check(lib().ga_tc_edge_support(self.ctx.handle, self._h, variant, _ptr(sup, _u32p), ctypes.byref(T)), "tc_edge_support")
        return T.value, sup[: self.m]

    # ---- peeling and core ----
    def batch_peel(self, num: int = 1, den: int = 10, T_full: Optional[int] = None, fused: bool = True) -> PeelResult:
        """Triangle-density batch peeling (densest.h:136-160 semantics)."""
        p = Peel(self)
        try:
            return p.run(num, den, T_full) if fused else p.run_rounds(num, den, T_full)
        finally:
            p.close()

    def greedy_pp(self, passes: int = 1):
        """Greedy++ for the triangle-density densest subgraph (passes = 1: Charikar's Greedy). Returns (tri, verts, set,
        per_pass, ms): the best density tri / verts found (0 / 0 without triangles), its vertices ascending, and the best
        (tri, verts) after every pass. The graph may have at most `ctx.greedy_max_vertices()` vertices."""
        passes = int(passes)
        t, v = ctypes.c_uint64(), ctypes.c_uint64()
        pp = np.zeros(2 * max(passes, 1), np.uint64)
        st = np.zeros(max(self.n, 1), np.uint32)
        cnt = ctypes.c_uint32()
        ms = ctypes.c_float()
        self.ctx.check(lib().ga_greedy_pp(self.ctx.handle, self._h, passes, ctypes.byref(t), ctypes.byref(v), _ptr(pp, _u64p),
                                          _ptr(st, _u32p), ctypes.byref(cnt), ctypes.byref(ms)), "greedy_pp")
        return t.value, v.value, st[: cnt.value], [(int(a), int(b)) for a, b in pp[: 2 * passes].reshape(-1, 2)], ms.value

    def truss_decomposition(self):
        """k-truss number of every oriented edge; returns (truss, rounds, kmax)."""
        t = Truss(self)
        try:
            return t.run()
        finally:
            t.close()

    def core_decomposition(self, incremental: bool = True, fused: bool = True, T_full: Optional[int] = None):
        """(k,triangle)-core numbers; returns (core, rounds, kmax)."""
        c = Core(self, incremental)
        try:
            if fused and incremental:
                return c.run(T_full)
            return c.run_rounds(T_full)
        finally:
            c.close()

    def close(self):
        if self._h:
            lib().ga_graph_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Truss:
    """k-truss decomposition state (ga_truss): edge-support peeling with the level rule of the core loops."""

    def __init__(self, g: DeviceGraph):
        self.g = g
        self.ctx = g.ctx
        self._h = _vp()
        self.ctx.check(lib().ga_truss_create(self.ctx.handle, g.handle, ctypes.byref(self._h)), "truss_create")
        self.last_ms = 0.0

    def run(self):
        """Returns (truss per oriented edge, rounds, kmax)."""
        stats = np.zeros(2, np.uint32)
        ms = ctypes.c_float()
        self.ctx.check(lib().ga_truss_run(self._h, _ptr(stats, _u32p), ctypes.byref(ms)), "truss_run")
        self.last_ms = ms.value
        t = np.zeros(max(self.g.m, 1), np.uint32)
        self.ctx.check(lib().ga_truss_download(self._h, _ptr(t, _u32p)), "truss_download")
        return t[: self.g.m], int(stats[0]), int(stats[1])

    def close(self):
        if self._h:
            lib().ga_truss_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Peel:
    """Batch-peeling state (ga_peel) with the reference's backend calls and the fused device loop."""

    def __init__(self, g: DeviceGraph):
        self.g = g
        self.ctx = g.ctx
        self._h = _vp()
        self.ctx.check(lib().ga_peel_create(self.ctx.handle, g.handle, ctypes.byref(self._h)), "peel_create")
        self.last_ms = 0.0

    def set_ctx(self, ctx: Context):
        """Runs this (idle) state on another context's stream from now on."""
        ctx.check(lib().ga_peel_set_ctx(self._h, ctx.handle), "peel_set_ctx")
        self.ctx = ctx

    # backend concept (densest.h:265-278)
    def reset_state(self):
        self.ctx.check(lib().ga_peel_reset_state(self._h), "peel_reset_state")

    def begin_round(self):
        self.ctx.check(lib().ga_peel_begin_round(self._h), "peel_begin_round")

    def run_tc(self):
        self.ctx.check(lib().ga_peel_run_tc(self._h), "peel_run_tc")

    def run_mark(self, nalive, num, den, rnd):
        self.ctx.check(lib().ga_peel_run_mark(self._h, nalive, num, den, rnd), "peel_run_mark")

    def read_stats(self):
        T, r = ctypes.c_uint64(), ctypes.c_uint64()
        self.ctx.check(lib().ga_peel_read_stats(self._h, ctypes.byref(T), ctypes.byref(r)), "peel_read_stats")
        return T.value, r.value

    def removed_round(self) -> np.ndarray:
        rr = np.zeros(max(self.g.n, 1), np.uint32)
        self.ctx.check(lib().ga_peel_download_removed_round(self._h, _ptr(rr, _u32p)), "download_removed_round")
        return rr[: self.g.n]

    def run(self, num=1, den=10, T_full: Optional[int] = None) -> PeelResult:
        stats = np.zeros(4, np.uint64)
        ms = ctypes.c_float()
        self.ctx.check(lib().ga_peel_run(self._h, (1 << 64) - 1 if T_full is None else T_full, num, den, _ptr(stats, _u64p),
                                         ctypes.byref(ms)), "peel_run")
        self.last_ms = ms.value
        return PeelResult(int(stats[0]), int(stats[1]), int(stats[2]), int(stats[3]), self.removed_round())

    def run_rounds(self, num=1, den=10, T_full: Optional[int] = None) -> PeelResult:
        """batch_peel_loop (densest.h:279-303) driven from the host, one readback per round."""
        n = self.g.n
        self.reset_state()
        nalive = n
        best_num = best_den = best_round = rounds = 0
        rnd = 0
        while nalive > 0:
            self.begin_round()
            self.run_tc()
            self.run_mark(nalive, num, den, rnd)
            T, removed = self.read_stats()
            if rnd == 0 and T_full is not None and T != T_full:
                raise GAError(GA_ERR_STATE, "device triangle count in round 0 differs from CPU count")
            if best_den == 0 or T * best_den > best_num * nalive:
                best_num, best_den, best_round = T, nalive, rnd
            if removed == 0 or removed > nalive:
                raise GAError(GA_ERR_STATE, "bad removed count from device (0 or more than alive)")
            nalive -= removed
            rounds = rnd + 1
            rnd += 1
        return PeelResult(best_num, best_den, best_round, rounds, self.removed_round())

    def close(self):
        if self._h:
            lib().ga_peel_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Core:
    """(k,triangle)-core state (ga_core): recount or incremental backend calls, fused incremental loop."""

    def __init__(self, g: DeviceGraph, incremental: bool = True):
        self.g = g
        self.ctx = g.ctx
        self.incremental = incremental
        self._h = _vp()
        self.ctx.check(lib().ga_core_create(self.ctx.handle, g.handle, 1 if incremental else 0, ctypes.byref(self._h)), "core_create")
        self.last_ms = 0.0

    def set_ctx(self, ctx: Context):
        """Runs this (idle) state on another context's stream from now on."""
        ctx.check(lib().ga_core_set_ctx(self._h, ctx.handle), "core_set_ctx")
        self.ctx = ctx

    def _stats(self):
        T, r, k = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        self.ctx.check(lib().ga_core_read_stats(self._h, ctypes.byref(T), ctypes.byref(r), ctypes.byref(k)), "core_read_stats")
        return T.value, r.value, k.value

    def core(self) -> np.ndarray:
        c = np.zeros(max(self.g.n, 1), np.uint32)
        self.ctx.check(lib().ga_core_download(self._h, _ptr(c, _u32p)), "core_download")
        return c[: self.g.n]

    def prune(self, k: int):
        """prune_to_core on the core numbers this (finished) state holds on the device: the subgraph on
        {v : core[v] >= k}; returns (DeviceGraph, old_id, core_out)."""
        n = self.g.n
        old = np.zeros(max(n, 1), np.uint32)
        cout = np.zeros(max(n, 1), np.uint32)
        h = _vp()
        ms = ctypes.c_float()
        self.ctx.check(lib().ga_core_prune(self._h, int(k), ctypes.byref(h), _ptr(old, _u32p), _ptr(cout, _u32p), ctypes.byref(ms)),
                       "core_prune")
        sub = DeviceGraph(self.ctx, h, ms.value)
        return sub, old[: sub.n], cout[: sub.n]

    def run(self, T_full: Optional[int] = None):
        stats = np.zeros(2, np.uint32)
        ms = ctypes.c_float()
        self.ctx.check(lib().ga_core_run(self._h, (1 << 64) - 1 if T_full is None else T_full, _ptr(stats, _u32p), ctypes.byref(ms)), "core_run")
        self.last_ms = ms.value
        return self.core(), int(stats[0]), int(stats[1])

    def run_rounds(self, T_full: Optional[int] = None):
        """core_decomp_loop / core_decomp_loop_inc (dsg.h:694-751) driven from the host."""
        L = lib()
        h = self._h
        self.ctx.check(L.ga_core_reset_state(h), "core_reset_state")
        n = self.g.n
        if self.incremental:
            T0 = ctypes.c_uint64()
            self.ctx.check(L.ga_core_init_tri(h, ctypes.byref(T0)), "core_init_tri")
            if T_full is not None and T0.value != T_full:
                raise GAError(GA_ERR_STATE, "device triangle count differs from CPU count")
        nalive, k, rounds, rnd = n, 0, 0, 0
        while nalive > 0:
            self.ctx.check(L.ga_core_begin_round(h), "core_begin_round")
            if self.incremental:
                self.ctx.check(L.ga_core_run_min(h), "run_min")
                self.ctx.check(L.ga_core_run_find(h, k), "run_find")
                self.ctx.check(L.ga_core_run_update(h), "run_update")
                self.ctx.check(L.ga_core_run_retire(h), "run_retire")
                _, removed, k_new = self._stats()
            else:
                self.ctx.check(L.ga_core_run_tc(h), "run_tc")
                self.ctx.check(L.ga_core_run_min(h), "run_min")
                self.ctx.check(L.ga_core_run_peel(h, k), "run_peel")
                T, removed, k_new = self._stats()
                if rnd == 0 and T_full is not None and T != T_full:
                    raise GAError(GA_ERR_STATE, "device triangle count in round 0 differs from CPU count")
            if removed == 0 or removed > nalive or k_new < k:
                raise GAError(GA_ERR_STATE, "bad stats from device (removed 0, removed > alive, or level decreased)")
            nalive -= removed
            k = k_new
            rounds = rnd + 1
            rnd += 1
        return self.core(), rounds, k

    def close(self):
        if self._h:
            lib().ga_core_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_tuning() -> TcTuning:
    """The library's default count tuning."""
    t = TcTuning()
    lib().ga_tc_default_tuning(ctypes.byref(t))
    return t


def variant_name(v: int) -> str:
    return lib().ga_tc_variant_name(v).decode()


def _bind_batch(L):
    if getattr(L, "_batch_bound", False):
        return
    sig = {
        "ga_batch_create": (ctypes.c_int, [_vp, ctypes.c_uint32, _u64p, _u32p, ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_float)]),
        "ga_batch_destroy": (None, [_vp]),
        "ga_batch_info": (ctypes.c_int, [_vp, _u32p, _u64p, _u64p]),
        "ga_batch_graph_info": (ctypes.c_int, [_vp, ctypes.c_uint32, _u32p, _u32p, _u32p]),
        "ga_batch_download_graph": (ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _u32p, _u32p, _u32p]),
        "ga_batch_count": (ctypes.c_int, [_vp, _vp, _u64p, ctypes.POINTER(ctypes.c_float)]),
        "ga_batch_count_per_graph": (ctypes.c_int, [_vp, _vp, _u64p, ctypes.POINTER(ctypes.c_float)]),
        "ga_batch_tune": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(TcTuning), ctypes.POINTER(ctypes.c_float)]),
        "ga_batch_set_tuning": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(TcTuning)]),
        "ga_batch_set_peel_graph": (ctypes.c_int, [_vp, ctypes.c_int]),
        "ga_batch_get_tuning": (ctypes.c_int, [_vp, ctypes.POINTER(TcTuning)]),
        "ga_batch_peel": (ctypes.c_int, [_vp, _vp, ctypes.c_uint64, ctypes.c_uint64, _u64p, _u64p, _u32p, ctypes.POINTER(ctypes.c_float)]),
        "ga_batch_peel_per_graph": (ctypes.c_int, [_vp, _vp, ctypes.c_uint64, ctypes.c_uint64, _u64p, _u64p, _u32p,
                                                   ctypes.POINTER(ctypes.c_float)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    L._batch_bound = True


def core_many(ctx: "Context", cores: Sequence["Core"], T_full: Optional[Sequence[Optional[int]]] = None):
    """Incremental core decompositions of several resident graphs in one launch (ga_core_run_many: one workgroup per
    graph, all rounds inside). Returns ([(core, rounds, kmax) per state], device ms)."""
    hs = (_vp * max(len(cores), 1))(*[c._h for c in cores])
    tf = None
    if T_full is not None:
        tf = np.array([(1 << 64) - 1 if t is None else int(t) for t in T_full], np.uint64)
    stats = np.zeros(max(2 * len(cores), 2), np.uint32)
    ms = ctypes.c_float()
    ctx.check(lib().ga_core_run_many(ctx.handle, hs, len(cores), None if tf is None else _ptr(tf, _u64p), _ptr(stats, _u32p),
                                     ctypes.byref(ms)), "core_run_many")
    return [(c.core(), int(stats[2 * i]), int(stats[2 * i + 1])) for i, c in enumerate(cores)], ms.value


def peel_batch(ctx: "Context", peels: Sequence["Peel"], num: int = 1, den: int = 10,
               T_full: Optional[Sequence[Optional[int]]] = None):
    """Fused peels of several resident graphs in shared launches (ga_peel_run_batch: every round's launches cover all
    unfinished states). The states are moved to ctx. Returns ([PeelResult per state], device ms)."""
    for p in peels:
        if p.ctx is not ctx:
            p.set_ctx(ctx)
    hs = (_vp * max(len(peels), 1))(*[p._h for p in peels])
    tf = None
    if T_full is not None:
        tf = np.array([(1 << 64) - 1 if t is None else int(t) for t in T_full], np.uint64)
    stats = np.zeros(max(4 * len(peels), 4), np.uint64)
    ms = ctypes.c_float()
    ctx.check(lib().ga_peel_run_batch(ctx.handle, hs, len(peels), None if tf is None else _ptr(tf, _u64p), num, den,
                                      _ptr(stats, _u64p), ctypes.byref(ms)), "peel_run_batch")
    out = [PeelResult(int(stats[4 * i]), int(stats[4 * i + 1]), int(stats[4 * i + 2]), int(stats[4 * i + 3]), p.removed_round())
           for i, p in enumerate(peels)]
    for p in peels:
        p.last_ms = ms.value
    return out, ms.value


def core_batch(ctx: "Context", cores: Sequence["Core"], T_full: Optional[Sequence[Optional[int]]] = None):
    """Fused incremental core decompositions of several resident graphs in shared launches (ga_core_run_batch: every
    round's three launches cover all unfinished states). The states are moved to ctx first. Returns
    ([(core, rounds, kmax) per state], device ms)."""
    for c in cores:
        if c.ctx is not ctx:
            c.set_ctx(ctx)
    hs = (_vp * max(len(cores), 1))(*[c._h for c in cores])
    tf = None
    if T_full is not None:
        tf = np.array([(1 << 64) - 1 if t is None else int(t) for t in T_full], np.uint64)
    stats = np.zeros(max(2 * len(cores), 2), np.uint32)
    ms = ctypes.c_float()
    ctx.check(lib().ga_core_run_batch(ctx.handle, hs, len(cores), None if tf is None else _ptr(tf, _u64p), _ptr(stats, _u32p),
                                      ctypes.byref(ms)), "core_run_batch")
    return [(c.core(), int(stats[2 * i]), int(stats[2 * i + 1])) for i, c in enumerate(cores)], ms.value


class Batch:
    """Many graphs resident together (ga_batch): per-graph counts and batch peeling with shared launches."""

    def __init__(self, ctx: Context, graphs: Sequence[np.ndarray] = (), flat: Optional[np.ndarray] = None,
                 offsets: Optional[np.ndarray] = None):
        """graphs: one raw pair array per graph; or flat (k, 2) uint32 pairs of all graphs with offsets (count + 1
        pair offsets), passed without a copy."""
        L = lib()
        _bind_batch(L)
        self.ctx = ctx
        if flat is None:
            sizes = [int(np.asarray(p).reshape(-1, 2).shape[0]) for p in graphs]
            off = np.zeros(len(graphs) + 1, np.uint64)
            off[1:] = np.cumsum(sizes, dtype=np.uint64)
            flat = np.concatenate([np.ascontiguousarray(p, dtype=np.uint32).reshape(-1) for p in graphs]) if graphs else np.zeros(0, np.uint32)
        else:
            off = np.ascontiguousarray(offsets, dtype=np.uint64)
        flat = np.ascontiguousarray(flat, dtype=np.uint32)
        count = len(off) - 1
        self._h = _vp()
        ms = ctypes.c_float()
        ctx.check(L.ga_batch_create(ctx.handle, count, _ptr(off, _u64p), _ptr(flat, _u32p) if flat.size else None,
                                    ctypes.byref(self._h), ctypes.byref(ms)), "batch_create")
        self.build_ms = ms.value
        c, tn, tm = ctypes.c_uint32(), ctypes.c_uint64(), ctypes.c_uint64()
        L.ga_batch_info(self._h, ctypes.byref(c), ctypes.byref(tn), ctypes.byref(tm))
        self.count, self.total_n, self.total_m = c.value, tn.value, tm.value
        self.n = np.zeros(self.count, np.int64)
        self.m = np.zeros(self.count, np.int64)
        for i in range(self.count):
            n, m, mo = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
            L.ga_batch_graph_info(self._h, i, ctypes.byref(n), ctypes.byref(m), ctypes.byref(mo))
            self.n[i], self.m[i] = n.value, m.value
        self.voff = np.concatenate([[0], np.cumsum(self.n)])

    def graph(self, i: int):
        rowptr = np.zeros(int(self.n[i]) + 1, np.uint32)
        col = np.zeros(max(int(self.m[i]), 1), np.uint32)
        src = np.zeros(max(int(self.m[i]), 1), np.uint32)
        self.ctx.check(lib().ga_batch_download_graph(self.ctx.handle, self._h, i, _ptr(rowptr, _u32p), _ptr(col, _u32p), _ptr(src, _u32p)),
                       "batch_download_graph")
        return rowptr, col[: self.m[i]], src[: self.m[i]]

    def count_all(self, per_graph: bool = False):
        T = np.zeros(self.count, np.uint64)
        ms = ctypes.c_float()
        f = lib().ga_batch_count_per_graph if per_graph else lib().ga_batch_count
        self.ctx.check(f(self.ctx.handle, self._h, _ptr(T, _u64p), ctypes.byref(ms)), "batch_count")
        return T, ms.value

    def work(self):
        """(sum over every graph's oriented edges of min(d+u, d+v), device ms of the last batched count)."""
        sm = ctypes.c_uint64()
        ms = ctypes.c_float()
        self.ctx.check(lib().ga_batch_work(self.ctx.handle, self._h, ctypes.byref(sm), ctypes.byref(ms)), "batch_work")
        return sm.value, ms.value

    def tune(self):
        """Autotunes the union's count; returns (tuning, best ms)."""
        t = TcTuning()
        ms = ctypes.c_float()
        self.ctx.check(lib().ga_batch_tune(self.ctx.handle, self._h, ctypes.byref(t), ctypes.byref(ms)), "batch_tune")
        return t, ms.value

    def get_tuning(self) -> TcTuning:
        t = TcTuning()
        lib().ga_batch_get_tuning(self._h, ctypes.byref(t))
        return t

    def set_tuning(self, t: TcTuning):
        """Count tuning of the union and of the per-graph baseline."""
        self.ctx.check(lib().ga_batch_set_tuning(self.ctx.handle, self._h, ctypes.byref(t)), "batch_set_tuning")

    def set_peel_graph(self, rounds: int):
        """Peel rounds captured into one HIP graph and replayed by peel_all (default 8; 0 enqueues every round)."""
        if lib().ga_batch_set_peel_graph(self._h, int(rounds)) != GA_OK:
            raise GAError(GA_ERR_ARG, f"peel graph rounds must be in [0, 64], got {rounds}")

    def peel_all(self, num=1, den=10, T_full=None, per_graph: bool = False, removed_round: bool = True):
        stats = np.zeros(4 * self.count, np.uint64)
        rr = np.zeros(max(int(self.total_n), 1), np.uint32) if removed_round else None
        tf = None if T_full is None else np.ascontiguousarray(T_full, dtype=np.uint64)
        ms = ctypes.c_float()
        f = lib().ga_batch_peel_per_graph if per_graph else lib().ga_batch_peel
        self.ctx.check(f(self.ctx.handle, self._h, num, den, None if tf is None else _ptr(tf, _u64p), _ptr(stats, _u64p),
                         None if rr is None else _ptr(rr, _u32p), ctypes.byref(ms)), "batch_peel")
        return stats.reshape(-1, 4), (rr[: self.total_n] if rr is not None else None), ms.value

    def close(self):
        if self._h:
            lib().ga_batch_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
