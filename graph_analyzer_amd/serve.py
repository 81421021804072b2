"""Serving loop: count / densest-subgraph (batch peel) / (k,triangle)-core requests on graphs kept resident in HBM.

No counterpart in the reference (it runs one graph per process); every request computes exactly what the single-graph
call computes (ga_tc_count_many / ga_peel_run / ga_core_run), so a result is the same alone and in any mix.

Shape on one MI355X:
  * graphs stay resident (oriented CSR + count plan; the undirected CSR once a core request needs it); a graph is
    evicted, least recently used first, only when it is idle and the footprint budget needs the room;
  * admission by device-memory footprint: a request runs when a worker slot is free and its graph plus the state it
    needs (a peel state: 9n bytes; a core state: undirected CSR + the fused loop's buffers) fit the budget;
  * continuous batching: `slots` workers, each with its own context (HIP stream), run admitted requests concurrently,
    so the kernels of different requests overlap on the device; requests are admitted the moment a slot and memory
    free up, not when a batch completes. The default of 4 workers matches the 4 hardware queues the HIP runtime gives a
    process: with more streams than queues, streams share a queue and a count waits behind the launches of a core
    batch on the same queue (measured: 4 workers served 335 / 443 / 559 req/s of the mixed load at 500 / 1,000 /
    2,000 offered, 16 workers 222 / 406 / 489);
  * stateful requests (peel, core) take one of up to `states_per_graph` state objects of their graph; states are
    created once and moved between worker streams (ga_peel_set_ctx / ga_core_set_ctx), so the steady state allocates
    and frees no device memory;
  * at most `core_slots` workers (default: half of them) run core work at a time, so counts and peels, which take
    milliseconds, do not wait behind core batches that take hundreds (same arrivals at 250 / 1,000 / 2,000 req/s,
    4 workers: p50 21.5 / 708 / 1,152 ms with every worker open to cores, 1.1 / 7.2 / 55 ms with two of them, at
    191-626 req/s against 203-617);
  * pending core requests (up to `core_batch` of them, on any graphs, each with its own state) are gathered into one
    ga_core_run_batch: every round's three launches cover all of them, so the round latency that bounds one
    decomposition is paid once per batch instead of once per request;
  * pending peel requests with the same eps (up to `peel_batch`, on any graphs) likewise go into one ga_peel_run_batch;
  * pending counts (up to `count_batch`) go into one ga_tc_count_many: one launch per bucket shape for all of them.
`slots=1, count_batch=1, core_batch=1, peel_batch=1` is the plain form: one request at a time, one count per launch,
one peel or decomposition at a time, in arrival order.
"""
from __future__ import annotations

import collections
import dataclasses
import hashlib
import itertools
import math
import queue
import threading
import time
from typing import Callable, Dict, List, Optional

import numpy as np

from . import TC_HASH, Context, Core, DeviceGraph, GAError, Peel, core_batch, peel_batch

KINDS = ("count", "peel", "core")


@dataclasses.dataclass
class Request:
    rid: int
    kind: str            # "count" | "peel" | "core"
    key: object          # graph key understood by the server's source
    arrival: float = 0.0  # seconds after the start of the run
    num: int = 1          # peel epsilon = num / den
    den: int = 20


@dataclasses.dataclass
class Result:
    rid: int
    kind: str
    key: object
    value: tuple         # count: (T,); peel: (best_num, best_den, best_round, rounds, removed_round digest);
    #                      core: (rounds, kmax, core digest)
    arrival: float
    start: float
    done: float
    worker: int
    batch: int           # requests that shared the launch sequence (count batching)

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def digest(a: np.ndarray) -> str:
    return hashlib.blake2b(np.ascontiguousarray(a).view(np.uint8), digest_size=16).hexdigest()


# device footprints (bytes), matching the library's allocations
def csr_bytes(n: int, m: int) -> int:
    return 4 * (n + 1) + 8 * m


def plan_bytes(m: int) -> int:
    return 20 * m + 4 * m + 4096          # work entries (5 words) + items (4 words each, about one per row)


def peel_bytes(n: int) -> int:
    return 9 * n + 4096


def core_bytes(n: int, m: int) -> int:
    m2 = max(2 * m, 1)
    nent = n + m2 // 8 + 16
    words = 8 * n + 2 * m2 + 2 * nent + 80            # alive lists, adjacency copies, ranges, frontier, cross lists, entries
    members = 4 * (5 * n + 24) + 64 + 16 * m2          # hub ids and table references; tables (at most m2 buckets)
    return 4 * (n + 1) + 4 * m2 + 13 * n + 4 * words + members + 4096


class _Resident:
    def __init__(self, key, dg: DeviceGraph, T: int, last: float):
        self.key = key
        self.dg = dg
        self.T = T
        self.n, self.m = dg.n, dg.m
        self.bytes = csr_bytes(self.n, self.m) + plan_bytes(self.m)
        self.idle: Dict[str, list] = {"peel": [], "core": []}
        self.nstates = {"peel": 0, "core": 0}
        self.inflight = 0
        self.last = last


class Server:
    def __init__(self, source: Callable[[object], np.ndarray], slots: int = 4, count_batch: int = 64,
                 states_per_graph: int = 4, budget_bytes: Optional[int] = None, device: int = 0,
                 core_batch: int = 1024, peel_batch: int = 1024, core_slots: Optional[int] = None):
        if slots < 1 or count_batch < 1 or states_per_graph < 1 or core_batch < 1 or peel_batch < 1:
            raise ValueError("slots, count_batch, core_batch, peel_batch and states_per_graph must be >= 1")
        if core_slots is None:   # half of the workers (measured best of 1-4 of 4: table in the design notes)
            core_slots = max(1, slots // 2)
        if not 1 <= core_slots <= slots:
            raise ValueError("core_slots must be between 1 and slots")
        self.source = source
        self.slots = slots
        self.core_slots = core_slots
        self.count_batch = count_batch
        self.core_batch = core_batch
        self.peel_batch = peel_batch
        self.states_per_graph = states_per_graph
        self.loader = Context(device)
        self.ctxs = [Context(device) for _ in range(slots)]
        free, total = self.loader.mem_info()
        self.budget = int(budget_bytes) if budget_bytes else int(free * 0.8)
        self.used = 0
        self.graphs: "collections.OrderedDict[object, _Resident]" = collections.OrderedDict()
        self.evictions = 0

    # ---- residency and footprint ---------------------------------------------------------------------------
    def _make_room(self, nbytes: int, keep=None) -> bool:
        while self.used + nbytes > self.budget:
            victim = next((r for r in self.graphs.values() if r.inflight == 0 and r is not keep), None)
            if victim is None:
                return False
            self._evict(victim)
        return True

    def _evict(self, r: _Resident):
        for kind in ("peel", "core"):
            for s in r.idle[kind]:
                s.close()
            r.idle[kind].clear()
        r.dg.close()
        self.used -= r.bytes
        del self.graphs[r.key]
        self.evictions += 1

    def load(self, key) -> Optional[_Resident]:
        """Makes a graph resident (CSR built on the GPU, count plan, T); None when the budget cannot hold it now."""
        r = self.graphs.get(key)
        if r is not None:
            self.graphs.move_to_end(key)
            return r
        pairs = self.source(key)
        est = csr_bytes(int(pairs.max()) + 1 if len(pairs) else 1, len(pairs)) + plan_bytes(len(pairs))
        if not self._make_room(est):
            return None
        dg = DeviceGraph.from_pairs(self.loader, pairs)
        dg.prepare()
        T, _ = dg.count(TC_HASH)
        r = _Resident(key, dg, int(T), time.perf_counter())
        self.graphs[key] = r
        self.used += r.bytes
        return r

    def preload(self, keys, states: bool = True):
        """Loads graphs (and one peel and one core state each) before serving, so serving itself allocates nothing."""
        for k in keys:
            r = self.load(k)
            if r is None:
                raise GAError(3, f"graph {k!r} does not fit the footprint budget")
            if states:
                for kind in ("peel", "core"):
                    s = self._new_state(r, kind)
                    if s is not None:
                        r.idle[kind].append(s)

    def _new_state(self, r: _Resident, kind: str):
        if r.nstates[kind] >= self.states_per_graph:
            return None
        nb = peel_bytes(r.n) if kind == "peel" else core_bytes(r.n, r.m)
        if not self._make_room(nb, keep=r):
            return None
        s = Peel(r.dg) if kind == "peel" else Core(r.dg, True)
        s._bytes = nb
        r.nstates[kind] += 1
        r.bytes += nb
        self.used += nb
        return s

    def _take_state(self, r: _Resident, kind: str):
        if r.idle[kind]:
            return r.idle[kind].pop()
        return self._new_state(r, kind)

    # ---- scheduling ----------------------------------------------------------------------------------------
    def _admit(self, pending: collections.deque, scan: int = 512, allow_core: bool = True):
        """Earliest admissible pending request (and, for a count, every other pending count up to count_batch). Core
        requests are skipped when allow_core is false (every worker that may run one is busy with one)."""
        # a snapshot: the deque is only changed right before returning, or restored to the same order
        for idx, req in enumerate(list(itertools.islice(pending, scan))):
            if req.kind == "core" and not allow_core:
                continue
            r = self.load(req.key)
            if r is None:
                continue
            if req.kind == "count":
                group, keep = [], collections.deque()
                while pending and len(group) < self.count_batch:
                    q = pending.popleft()
                    rq = self.load(q.key) if q.kind == "count" else None
                    if rq is not None:
                        group.append((q, rq))
                    else:
                        keep.append(q)
                pending.extendleft(reversed(keep))
                for _, rq in group:
                    rq.inflight += 1
                return ("count", group, None)
            if req.kind == "peel" and self.peel_batch > 1:
                # every pending peel request with this eps whose graph is resident and has a free state goes into one
                # ga_peel_run_batch (shared launches for all of their rounds)
                group, states, keep = [], [], collections.deque()
                while pending and len(group) < self.peel_batch:
                    q = pending.popleft()
                    same = q.kind == "peel" and (q.num, q.den) == (req.num, req.den)
                    rq = self.graphs.get(q.key) if same else None
                    st = self._take_state(rq, "peel") if rq is not None else None
                    if st is not None:
                        group.append((q, rq))
                        states.append(st)
                        rq.inflight += 1
                    else:
                        keep.append(q)
                pending.extendleft(reversed(keep))
                if group:
                    return ("peel_batch", group, states)
                continue
            if req.kind == "core" and self.core_batch > 1:
                # every pending core request whose graph is resident and has a free state goes into one
                # ga_core_run_batch (shared launches for all of their rounds)
                group, states, keep = [], [], collections.deque()
                while pending and len(group) < self.core_batch:
                    q = pending.popleft()
                    rq = self.graphs.get(q.key) if q.kind == "core" else None
                    st = self._take_state(rq, "core") if rq is not None else None
                    if st is not None:
                        group.append((q, rq))
                        states.append(st)
                        rq.inflight += 1
                    else:
                        keep.append(q)
                pending.extendleft(reversed(keep))
                if group:
                    return ("core_batch", group, states)
                continue
            state = self._take_state(r, req.kind)
            if state is None:
                continue
            del pending[idx]
            r.inflight += 1
            return (req.kind, [(req, r)], state)
        return None

    def _execute(self, w: int, task):
        kind, group, state = task
        ctx = self.ctxs[w]
        start = time.perf_counter()
        if kind == "count":
            T, _ = ctx.count_many([r.dg for _, r in group], TC_HASH)
            vals = [(int(t),) for t in T]
        elif kind == "core_batch":
            res, _ = core_batch(ctx, state, [r.T for _, r in group])
            vals = [(rounds, kmax, digest(core)) for core, rounds, kmax in res]
        elif kind == "peel_batch":
            req0 = group[0][0]
            res, _ = peel_batch(ctx, state, req0.num, req0.den, [r.T for _, r in group])
            vals = [(pr.best_num, pr.best_den, pr.best_round, pr.rounds, digest(pr.removed_round)) for pr in res]
        elif kind == "peel":
            (req, r), = group
            state.set_ctx(ctx)
            pr = state.run(req.num, req.den, T_full=r.T)
            vals = [(pr.best_num, pr.best_den, pr.best_round, pr.rounds, digest(pr.removed_round))]
        else:
            (req, r), = group
            state.set_ctx(ctx)
            core, rounds, kmax = state.run(T_full=r.T)
            vals = [(rounds, kmax, digest(core))]
        return start, time.perf_counter(), vals

    def run(self, requests: List[Request], realtime: bool = True) -> List[Result]:
        """Serves `requests` (arrivals replayed in real time when realtime) and returns one Result per request."""
        reqs = sorted(requests, key=lambda q: (q.arrival, q.rid))
        todo: List[queue.Queue] = [queue.Queue() for _ in range(self.slots)]
        done: queue.Queue = queue.Queue()

        def worker(w):
            while True:
                task = todo[w].get()
                if task is None:
                    return
                try:
                    done.put((w, task, self._execute(w, task), None))
                except Exception as e:   # reported by the scheduler thread
                    done.put((w, task, None, e))

        threads = [threading.Thread(target=worker, args=(w,), daemon=True) for w in range(self.slots)]
        for t in threads:
            t.start()
        pending: collections.deque = collections.deque()
        free = list(range(self.slots - 1, -1, -1))
        core_busy = 0   # workers running a core task: at most core_slots, so counts and peels never wait behind cores
        out: Dict[int, Result] = {}
        error = None
        t0 = time.perf_counter()
        i = 0
        try:
            while len(out) < len(reqs) and error is None:
                now = time.perf_counter() - t0
                while i < len(reqs) and (not realtime or reqs[i].arrival <= now):
                    pending.append(reqs[i])
                    i += 1
                refused = False   # the last admission attempt found nothing it could start
                while free and pending:
                    task = self._admit(pending, allow_core=core_busy < self.core_slots)
                    if task is None:
                        refused = True
                        break
                    if task[0] in ("core", "core_batch"):
                        core_busy += 1
                    todo[free.pop()].put(task)
                if refused and len(free) == self.slots and i >= len(reqs):
                    # nothing in flight, nothing left to arrive, and the pending requests cannot start
                    raise GAError(3, f"{len(pending)} requests cannot be admitted: their graphs do not fit the budget")
                wait = 0.02 if i >= len(reqs) else max(0.0, min(0.02, reqs[i].arrival - (time.perf_counter() - t0)))
                try:
                    item = done.get(timeout=wait) if wait > 0 else done.get_nowait()
                except queue.Empty:
                    continue
                while item is not None:
                    w, task, res, err = item
                    free.append(w)
                    kind, group, state = task
                    if kind in ("core", "core_batch"):
                        core_busy -= 1
                    for _, r in group:
                        r.inflight -= 1
                        r.last = time.perf_counter()
                    if kind in ("core_batch", "peel_batch"):
                        for (_, r), st in zip(group, state):
                            r.idle[kind[:4]].append(st)
                    elif state is not None:
                        group[0][1].idle[kind].append(state)
                    if err is not None:
                        error = err
                        break
                    start, end, vals = res
                    for (req, _), v in zip(group, vals):
                        out[req.rid] = Result(req.rid, req.kind, req.key, v, req.arrival, start - t0, end - t0, w, len(group))
                    try:
                        item = done.get_nowait()
                    except queue.Empty:
                        item = None
        finally:
            for q in todo:
                q.put(None)
            for t in threads:
                t.join()
        if error is not None:
            raise error
        return [out[q.rid] for q in reqs]

    def close(self):
        for r in list(self.graphs.values()):
            self._evict(r)
        for c in self.ctxs:
            c.close()
        self.loader.close()


def summarize(results: List[Result]) -> dict:
    lat = np.array([r.latency for r in results]) * 1e3
    span = max(r.done for r in results) - min(r.arrival for r in results)
    by = {}
    for kind in KINDS:
        k = np.array([r.latency for r in results if r.kind == kind]) * 1e3
        if k.size:
            by[kind] = {"n": int(k.size), "p50_ms": float(np.percentile(k, 50)), "p99_ms": float(np.percentile(k, 99))}
    return {"requests": len(results), "throughput_rps": len(results) / span, "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "max_ms": float(lat.max()), "makespan_s": span, "by_kind": by,
            "mean_count_batch": float(np.mean([r.batch for r in results if r.kind == "count"] or [0]))}


def poisson_requests(rate: float, duration: float, seed: int = 2106, pool: int = 256, rmat_graphs: int = 2,
                     rmat_share: float = 0.05, num: int = 1, den: int = 20) -> List[Request]:
    """Poisson arrivals at `rate` req/s for `duration` s, uniform over the kinds; a request picks one of `pool` graphs
    ("mix", i) with probability 1 - rmat_share, else one of `rmat_graphs` R-MAT graphs ("rmat", j)."""
    rng = np.random.default_rng(seed)
    t, reqs = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        kind = KINDS[int(rng.integers(3))]
        key = ("rmat", int(rng.integers(rmat_graphs))) if rng.random() < rmat_share else ("mix", int(rng.integers(pool)))
        reqs.append(Request(len(reqs), kind, key, t, num, den))
    return reqs


def workload_source(seed: int = 2106, pool: int = 256, rmat_scale: int = 20, nmax: int = 200_000):
    """The graphs requests name: ("mix", i) is graph i of `pool` graphs of the batch size law, ("rmat", j) an R-MAT
    graph of scale `rmat_scale`, edge factor 8. Each is generated once, on first use."""
    from . import gen
    specs = gen.batch_mix_specs(pool, seed=seed, nmax=nmax)
    cache: Dict[object, np.ndarray] = {}

    def source(key):
        if key not in cache:
            if key[0] == "mix":
                cache[key] = gen.batch_graph(specs[key[1]])
            else:
                cache[key] = gen.rmat(rmat_scale, 8 << rmat_scale, seed=seed * 10 + key[1])
        return cache[key]

    return source


def mixed_workload(rate: float, duration: float, seed: int = 2106, pool: int = 256, rmat_graphs: int = 2,
                   rmat_scale: int = 20, rmat_share: float = 0.05, num: int = 1, den: int = 20, nmax: int = 200_000):
    """Poisson arrivals at `rate` req/s for `duration` s of count / peel (eps num/den) / core requests, uniform over the
    kinds; a request picks one of `pool` graphs of the batch size law (95%) or one of `rmat_graphs` R-MAT graphs of
    scale `rmat_scale`, edge factor 8 (5%). Returns (source, requests)."""
    source = workload_source(seed, pool, rmat_scale, nmax)
    return source, poisson_requests(rate, duration, seed, pool, rmat_graphs, rmat_share, num, den)
