"""GPU parity of induced_subgraph / prune_to_core (dsg.h:103-114, 636-645): the device-built subgraph CSR, old ids and
restricted core numbers equal the reference's fixtures and the CPU oracle bit for bit, from host core numbers
(ga_graph_prune_to_core), from a finished device decomposition (ga_core_prune), and for arbitrary keep masks
(ga_graph_induced), on fixtures, random graphs with duplicates and loops, power-law graphs and empty selections."""
import ctypes
import json
import os

import numpy as np
import pytest

import graph_analyzer_amd as ga
import oracle
from conftest import random_graph

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")


def same_graph(dg, og, tag):
    r, c, s = dg.download()
    orr, oc, os_ = og.arrays()
    assert (dg.n, dg.m, dg.max_out_degree) == (og.n, og.m, og.max_out_degree), tag
    assert np.array_equal(r, orr) and np.array_equal(c, oc) and np.array_equal(s, os_), f"subgraph CSR differs {tag}"


def test_prune_golden(ctx):
    """Every fixture graph at k in {0, 1, ceil(kmax/2), kmax, kmax+1}: host core numbers and device ones."""
    with open(GOLDEN) as f:
        graphs = json.load(f)["graphs"]
    for rec in graphs:
        raw = np.array(rec["raw_pairs"], dtype=np.uint32).reshape(-1, 2)
        g = ga.DeviceGraph.from_pairs(ctx, raw)
        core = np.array(rec["core"], dtype=np.uint32)
        c = ga.Core(g, True)
        cg, _, kmax = c.run(T_full=rec["T"])
        assert np.array_equal(cg, core) and kmax == rec["kmax"], rec["name"]
        for p in rec["prune"]:
            for how in ("host", "device"):
                sub, old, cout = g.prune_to_core(core, p["k"]) if how == "host" else c.prune(p["k"])
                r, col, src = sub.download()
                tag = (rec["name"], p["k"], how)
                assert (sub.n, sub.m, sub.max_out_degree) == (p["n"], p["m"], p["max_out_degree"]), tag
                assert r.tolist() == p["rowptr"] and col.tolist() == p["col"], tag
                assert np.array_equal(src, np.repeat(np.arange(sub.n, dtype=np.uint32), np.diff(r.astype(np.int64)))), tag
                assert old.tolist() == p["old_id"] and cout.tolist() == p["core_out"], tag
                sub.close()
        c.close()
        g.close()


@pytest.mark.parametrize("seed", range(4))
def test_induced_random_masks(ctx, seed):
    """Random keep masks (all, none, and fractions in between) on random graphs with duplicates and self loops."""
    rng = np.random.default_rng(300 + seed)
    n = int(rng.integers(50, 3000))
    raw = random_graph(rng, n, float(rng.uniform(0.002, 0.05)), dup=0.3, loops=3)
    og = oracle.OGraph.from_pairs(raw)
    g = ga.DeviceGraph.from_pairs(ctx, raw)
    for p in (1.0, 0.95, 0.6, 0.25, 0.01, 0.0):
        keep = (rng.random(g.n) < p).astype(np.uint8)
        keep[keep.astype(bool)] = rng.integers(1, 256, int(keep.sum()), dtype=np.uint8)   # any non-zero byte keeps
        sub, old = g.induced(keep)
        osub, oold = og.induced(keep)
        assert np.array_equal(old, oold), (seed, p)
        same_graph(sub, osub, (seed, p))
        if sub.m:
            assert sub.count(ga.TC_HASH)[0] == osub.triangles() == og.participation((keep != 0).astype(np.uint8))[0]
        sub.close()
    g.close()


def test_prune_pipeline_powerlaw(ctx):
    """The gpu_dsg order (gpu_dsg.cu:274-328) on the 8,298-vertex power-law workload and two Chung-Lu graphs: device
    core numbers, the kmax-core density as lower bound, prune at ceil(lb) on the device, then the pruned graph's own
    device core numbers must equal the restricted ones."""
    from graph_analyzer_amd import gen
    raws = [gen.powerlaw_small(), gen.chung_lu(20000, 150000, 2.1, 0.0, seed=41), gen.chung_lu(60000, 500000, 2.3, 5.0, seed=42)]
    for i, raw in enumerate(raws):
        og = oracle.OGraph.from_pairs(raw)
        T = og.triangles_omp(0)
        g = ga.DeviceGraph.from_pairs(ctx, raw)
        c = ga.Core(g, True)
        core, _, kmax = c.run(T_full=T)
        ocore, okmax = og.core()
        assert np.array_equal(core, ocore) and kmax == okmax
        kt, kv = og.core_level_density(ocore, okmax)
        for k in sorted({(kt + kv - 1) // kv, max(kmax // 4, 1), kmax}):
            sub, old, cout = c.prune(k)
            osub, oold, ocout = og.prune_to_core(ocore, k)
            assert np.array_equal(old, oold) and np.array_equal(cout, ocout), (i, k)
            same_graph(sub, osub, (i, k))
            if sub.n:
                Ts = sub.count(ga.TC_HASH)[0]
                own, _, _ = sub.core_decomposition(T_full=Ts)
                assert np.array_equal(own, cout), (i, k)
            sub.close()
        c.close()
        g.close()


def test_subgraph_errors(ctx):
    g = ga.DeviceGraph.from_pairs(ctx, np.array([[0, 1], [1, 2], [0, 2]], np.uint32))
    with pytest.raises(ValueError):
        g.induced(np.ones(2, np.uint8))
    out = ga._vp()
    rc = ga.lib().ga_graph_prune_to_core(ctx.handle, g.handle, None, 1, ctypes.byref(out), None, None, None)
    assert rc == ga.GA_ERR_ARG and not out
    rc = ga.lib().ga_graph_induced(ctx.handle, g.handle, None, ctypes.byref(out), None, None)
    assert rc == ga.GA_ERR_ARG and not out
    sub, old = g.induced(np.zeros(3, np.uint8))    # nothing kept: a graph with no vertices
    assert (sub.n, sub.m, old.size) == (0, 0, 0)
    sub.close()
    g.close()
