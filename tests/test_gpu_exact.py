"""CoreExact (dsg.h:454-627) and Exact (dsg.h:402-428) through ga_core_exact / ga_exact_densest, against the reference's
own results (tests/golden/exact_vectors.json, written by the reference's code through tests/golden/make_exact_golden.py)
and brute force: the best density, the vertex set, and the search's counters (flow calls, components searched, network
shrinks) must be identical, with Pruning1/2, without them (the reference's test hook, which makes the network shrink over
many core levels) and seeded with Greedy++ as gpu_dsg.cu:348-361 runs it."""
import json
import os

import numpy as np
import pytest

import graph_analyzer_amd as ga
import oracle
from conftest import clique, random_graph

pytestmark = pytest.mark.gpu

FIXTURES = os.path.join(os.path.dirname(__file__), "golden", "exact_vectors.json")
KEYS = ("tri", "verts", "set", "flow_calls", "components_searched", "shrinks")


def fixtures():
    with open(FIXTURES) as f:
        return json.load(f)["graphs"]


def same(got, exp, tag):
    for k in KEYS:
        assert got[k] == exp[k], (tag, k, got[k], exp[k])


def test_core_exact_fixtures(ctx):
    for rec in fixtures():
        raw = np.array(rec["raw_pairs"], np.uint32).reshape(-1, 2)
        g = ga.DeviceGraph.from_pairs(ctx, raw)
        assert (g.n, g.m) == (rec["n"], rec["m"]), rec["name"]
        core = np.array(rec["core"], np.uint32)
        same(g.core_exact(core), rec["core_exact"], (rec["name"], "pruning12"))
        same(g.core_exact(core, pruning12=False), rec["core_exact_no_pruning12"], (rec["name"], "no pruning12"))
        gp = rec["greedy_pp10"]
        seed = (gp["tri"], gp["verts"], gp["set"]) if gp["verts"] else None
        same(g.core_exact(core, seed=seed), rec["core_exact_seeded"], (rec["name"], "seeded"))
        if "exact_densest" in rec:
            same(g.exact_densest(), rec["exact_densest"], (rec["name"], "exact_densest"))
        g.close()


def test_core_exact_on_device_core_numbers_and_greedy_seed(ctx):
    """The pipeline of gpu_dsg.cu:318-361 on the device: core numbers from the fused level loop, Greedy++ on the graph,
    CoreExact seeded with it. Equal to the fixture's (reference) result."""
    for rec in fixtures():
        if rec["kmax"] < 3:
            continue
        raw = np.array(rec["raw_pairs"], np.uint32).reshape(-1, 2)
        g = ga.DeviceGraph.from_pairs(ctx, raw)
        c = ga.Core(g, True)
        core, _, kmax = c.run()
        c.close()
        assert kmax == rec["kmax"] and np.array_equal(core, np.array(rec["core"], np.uint32)), rec["name"]
        t, v, s, _, _ = g.greedy_pp(10)
        got = g.core_exact(core, seed=(t, v, s))
        same(got, rec["core_exact_seeded"], rec["name"])
        g.close()


def test_exact_matches_brute_force(ctx):
    """Small random graphs: the best density of both searches equals the exhaustive maximum over all vertex subsets,
    and the returned set really has that many triangles."""
    rng = np.random.default_rng(4242)
    for i in range(24):
        n = int(rng.integers(6, 16))
        raw = random_graph(rng, n, float(rng.uniform(0.3, 0.8)))
        og = oracle.OGraph.from_pairs(raw)
        if og.triangles() == 0:
            continue
        r, c, s = og.arrays()
        uniq = np.stack([s, c], 1)
        bn, bd = oracle.brute_densest(og.n, uniq)
        g = ga.DeviceGraph.from_pairs(ctx, raw)
        core, _ = og.core()
        for res in (g.core_exact(core), g.core_exact(core, pruning12=False), g.exact_densest()):
            assert res["tri"] * bd == bn * res["verts"], (i, res, (bn, bd))
            keep = np.zeros(og.n, np.uint8)
            keep[res["set"]] = 1
            assert og.participation(keep)[0] == res["tri"] and len(res["set"]) == res["verts"], i
        g.close()


def test_core_exact_errors(ctx):
    """The reference's checks: core numbers of the wrong length, an invalid seed (wrong triangle count, a duplicate,
    an id out of range), core numbers inconsistent with the graph."""
    raw = np.array(clique(6) + clique(4, base=6), np.uint32)
    g = ga.DeviceGraph.from_pairs(ctx, raw)
    og = oracle.OGraph.from_pairs(raw)
    core, _ = og.core()
    ok = g.core_exact(core)
    assert (ok["tri"], ok["verts"]) == (20, 6)
    with pytest.raises(ValueError):
        g.core_exact(core[:-1])
    for seed in ((5, 3, [0, 1, 2]), (1, 3, [0, 0, 1]), (1, 3, [0, 1, 99])):
        with pytest.raises(ga.GAError):
            g.core_exact(core, seed=seed)
    with pytest.raises(ga.GAError):
        g.core_exact(np.zeros(og.n, np.uint32) + 1)   # kmax 1 but the graph holds a K6 (density 20/6 > 1)
    g.close()
