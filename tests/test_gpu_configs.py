"""The batching and serving configurations at their full size, checked against the CPU oracle.

* The 4,096-graph mix (n log-uniform in [1k, 200k], mean degree uniform in [4, 64], half Erdos-Renyi and half Chung-Lu,
  about 2.6B edges) resident as one disjoint union: the batched counts equal one count per graph on every graph, and
  the batched eps = 1/10 peel equals one fused peel per graph (every statistic and every vertex's removal round); 24
  graphs spread over the size range are rebuilt from their raw pairs by the oracle and compared on the CSR bytes, T,
  the peel's 4-tuple and the removal rounds.
* Serving at 2,000 req/s for 1 s (count / peel at eps = 1/20 / core, 95% on 256 graphs of the batch size law and 5% on
  two R-MAT scale-20 graphs), continuous batching on 16 worker streams: every one of the ~2,000 results equals the
  oracle's result for its (kind, graph), so each result is the same as the single-graph computation alone.

The oracle work runs on a thread pool (its calls release the GIL), one graph per thread."""
import concurrent.futures as cf
import os

import numpy as np
import pytest

import graph_analyzer_amd as ga
import oracle
from graph_analyzer_amd import gen, serve

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

WORKERS = min(16, os.cpu_count() or 4)


def test_batch_mix_4096_full_size(ctx):
    specs = gen.batch_mix_specs(4096)
    bp = gen.BatchPairs(specs)
    b = ga.Batch(ctx, flat=bp.pairs, offsets=bp.offsets)
    print(f"batch: {b.count} graphs, n {b.total_n}, m {b.total_m}", flush=True)
    assert b.count == 4096 and b.total_m > 2_000_000_000
    T, _ = b.count_all()
    Tp, _ = b.count_all(per_graph=True)
    assert np.array_equal(T, Tp), "batched counts differ from one count per graph"
    st, rr, _ = b.peel_all(1, 10, T_full=T)
    st2, rr2, _ = b.peel_all(1, 10, T_full=T, per_graph=True)
    assert np.array_equal(st, st2) and np.array_equal(rr, rr2), "batched peel differs from one peel per graph"
    print(f"batch: T total {int(T.sum())}, counts and peels equal per graph", flush=True)
    order = np.argsort(b.m, kind="stable")
    sample = sorted(set(int(order[k]) for k in np.linspace(0, b.count - 1, 24).astype(int)))
    off = bp.offsets

    def check(i):
        raw = bp.pairs[int(off[i]): int(off[i + 1])]
        og = oracle.OGraph.from_pairs(raw)
        r, c, s = b.graph(i)
        for x, y in zip((r, c, s), og.arrays()):
            if not np.array_equal(x, y):
                return f"graph {i}: CSR differs"
        if og.triangles() != int(T[i]):
            return f"graph {i}: T differs"
        bn, bd, br, rounds, orr = og.batch_peel(1, 10)
        if tuple(int(x) for x in st[i]) != (bn, bd, br, rounds):
            return f"graph {i}: peel statistics differ"
        v0 = int(b.voff[i])
        if not np.array_equal(rr[v0: v0 + int(b.n[i])], orr):
            return f"graph {i}: removal rounds differ"
        return None

    with cf.ThreadPoolExecutor(WORKERS) as ex:
        errors = [e for e in ex.map(check, sample) if e]
    assert not errors, errors
    print(f"batch: {len(sample)} graphs (m {int(b.m[sample].min())}..{int(b.m[sample].max())}) equal the oracle", flush=True)
    b.close()
    del bp


def test_serving_2000rps_full_size():
    source, reqs = serve.mixed_workload(rate=2000.0, duration=1.0, seed=2106, pool=256, rmat_scale=20)
    keys = sorted({r.key for r in reqs})
    assert len(reqs) > 1800 and {r.kind for r in reqs} == set(serve.KINDS) and any(k[0] == "rmat" for k in keys)
    s = serve.Server(source, slots=16, count_batch=64, states_per_graph=4)
    s.preload(keys)
    got = s.run(reqs)
    s.close()
    summ = serve.summarize(got)
    print(f"serving: {summ['requests']} requests, {summ['throughput_rps']:.0f} req/s, p50 {summ['p50_ms']:.1f} ms, "
          f"p99 {summ['p99_ms']:.1f} ms", flush=True)
    need = sorted({(r.kind, r.key) for r in got})

    def expected(item):
        kind, key = item
        og = oracle.OGraph.from_pairs(source(key))
        if kind == "count":
            return item, (og.triangles(),)
        if kind == "peel":
            bn, bd, br, rounds, rr = og.batch_peel(1, 20)
            return item, (bn, bd, br, rounds, serve.digest(rr.astype(np.uint32)))
        core, kmax = og.core()
        _, rounds, _ = og.core_levels()
        return item, (rounds, kmax, serve.digest(core.astype(np.uint32)))

    with cf.ThreadPoolExecutor(WORKERS) as ex:
        want = dict(ex.map(expected, need))
    bad = [(r.rid, r.kind, r.key) for r in got if r.value != want[(r.kind, r.key)]]
    assert not bad, bad[:10]
    print(f"serving: all {len(got)} results equal the oracle ({len(need)} distinct (kind, graph))", flush=True)
