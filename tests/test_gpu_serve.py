"""Serving loop: every result is the same alone and in the mix (bit-exact), and equals the CPU oracle.

The mix: Poisson arrivals of count / peel (eps 1/20) / core requests over a pool of small graphs of the batch size law
plus R-MAT graphs, served by 8 concurrent worker streams with count batching; "alone" replays every request by itself
(one worker, one graph per count launch, no overlap)."""
import numpy as np
import pytest

import graph_analyzer_amd as ga
import oracle
from graph_analyzer_amd import serve

pytestmark = pytest.mark.gpu


def _workload():
    return serve.mixed_workload(rate=600.0, duration=0.5, seed=7, pool=24, rmat_graphs=2, rmat_scale=12, rmat_share=0.1,
                                nmax=20_000)


def test_alone_equals_mix_and_oracle():
    source, reqs = _workload()
    assert len(reqs) > 150 and {r.kind for r in reqs} == set(serve.KINDS)
    keys = sorted({r.key for r in reqs})
    mix = serve.Server(source, slots=8, count_batch=16, states_per_graph=2)
    mix.preload(keys)
    got = mix.run(reqs)
    mix.close()
    alone = serve.Server(source, slots=1, count_batch=1, core_batch=1, peel_batch=1, states_per_graph=1)
    ref = alone.run(reqs, realtime=False)
    alone.close()
    assert [r.rid for r in got] == [r.rid for r in ref]
    for a, b in zip(got, ref):
        assert a.value == b.value, (a.rid, a.kind, a.key)
    assert max(r.batch for r in got if r.kind == "count") > 1, "no count requests were batched"
    assert max(r.batch for r in got if r.kind == "core") > 1, "no core requests shared a batch"
    # the CPU oracle on one request of each kind per graph
    seen = set()
    for r in got:
        if (r.kind, r.key) in seen:
            continue
        seen.add((r.kind, r.key))
        og = oracle.OGraph.from_pairs(source(r.key))
        if r.kind == "count":
            assert r.value == (og.triangles(),)
        elif r.kind == "peel":
            bn, bd, br, rounds, rr = og.batch_peel(1, 20)
            assert r.value == (bn, bd, br, rounds, serve.digest(rr.astype(np.uint32)))
        else:
            core, kmax = og.core()
            _, rounds, _ = og.core_levels()
            assert r.value == (rounds, kmax, serve.digest(core.astype(np.uint32)))


def test_footprint_budget_evicts_idle_graphs():
    source, reqs = _workload()
    reqs = reqs[:60]
    keys = sorted({r.key for r in reqs})
    sizes = []
    for k in keys:
        p = source(k)
        sizes.append(serve.csr_bytes(int(p.max()) + 1, len(p)) + serve.plan_bytes(len(p)) + serve.peel_bytes(int(p.max()) + 1)
                     + serve.core_bytes(int(p.max()) + 1, len(p)))
    budget = 2 * max(sizes)   # room for about two graphs with their states: the rest must be evicted and reloaded
    s = serve.Server(source, slots=2, count_batch=4, states_per_graph=1, budget_bytes=budget)
    got = s.run(reqs)
    assert s.evictions > 0 and s.used <= budget
    s.close()
    alone = serve.Server(source, slots=1, count_batch=1, core_batch=1, peel_batch=1, states_per_graph=1)
    ref = alone.run(reqs, realtime=False)
    alone.close()
    assert [a.value for a in got] == [b.value for b in ref]


def test_states_created_while_serving():
    """Core and peel states created on demand while other requests run on the worker streams (creation must be
    finished before a state moves to a worker's stream): many stateful requests on few graphs, nothing preloaded."""
    source, reqs = serve.mixed_workload(rate=2000.0, duration=0.2, seed=11, pool=4, rmat_graphs=1, rmat_scale=12,
                                        rmat_share=0.1, nmax=8_000)
    s = serve.Server(source, slots=8, count_batch=8, states_per_graph=4)
    got = s.run(reqs)
    assert max(r.nstates["core"] for r in s.graphs.values()) > 1
    s.close()
    alone = serve.Server(source, slots=1, count_batch=1, core_batch=1, peel_batch=1, states_per_graph=1)
    ref = alone.run(reqs, realtime=False)
    alone.close()
    assert [a.value for a in got] == [b.value for b in ref]
