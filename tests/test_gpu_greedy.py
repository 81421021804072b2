"""GPU parity of Greedy++ (ga_greedy_pp, dsg.h:189-250): the best density, its vertex set and the best after every pass
equal the reference's on the committed fixtures and the CPU oracle's (itself pinned against the reference build, the
exhaustive optimum and closed forms in tests/test_oracle.py) on random graphs around the bitmap word size, power-law
graphs, pruned graphs of the core pipeline, the largest graph one workgroup takes, and degenerate inputs."""
import json
import math
import os

import numpy as np
import pytest

import graph_analyzer_amd as ga
import oracle
from conftest import clique, random_graph
from graph_analyzer_amd import gen

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")


def same(got, exp, tag):
    t, v, s, pp, _ = got
    et, ev, es, epp = exp
    assert (t, v) == (et, ev), tag
    assert np.array_equal(s, np.asarray(es, np.uint32)), tag
    assert [tuple(x) for x in pp] == [tuple(x) for x in epp], tag


def check(ctx, raw, passes, tag):
    og = oracle.OGraph.from_pairs(raw)
    g = ga.DeviceGraph.from_pairs(ctx, raw)
    for p in passes:
        exp = og.greedy_pp(p)
        same(g.greedy_pp(p), exp, (tag, p))
    g.close()


def test_greedy_fixtures(ctx):
    print(f"greedy_pp takes up to {ctx.greedy_max_vertices()} vertices", flush=True)
    cases = 0
    with open(GOLDEN) as f:
        for rec in json.load(f)["graphs"]:
            raw = np.array(rec["raw_pairs"], dtype=np.uint32).reshape(-1, 2)
            g = ga.DeviceGraph.from_pairs(ctx, raw)
            for passes, e in rec["greedy_pp"].items():
                same(g.greedy_pp(int(passes)), (e["tri"], e["verts"], e["set"], e["per_pass"]), (rec["name"], passes))
                cases += 1
            g.close()
    assert cases >= 90


def test_greedy_closed_forms(ctx):
    for k in (3, 4, 9, 40):
        g = ga.DeviceGraph.from_pairs(ctx, np.array(clique(k), np.uint32))
        t, v, s, pp, _ = g.greedy_pp(3)
        assert (t, v, s.tolist(), pp) == (math.comb(k, 3), k, list(range(k)), [(math.comb(k, 3), k)] * 3)
        g.close()
    g = ga.DeviceGraph.from_pairs(ctx, np.array(clique(4) + clique(5, 4), np.uint32))
    assert g.greedy_pp(2)[:2] == (10, 5) and g.greedy_pp(2)[2].tolist() == [4, 5, 6, 7, 8]
    g.close()
    star = np.array([(0, i) for i in range(1, 70)], np.uint32)
    g = ga.DeviceGraph.from_pairs(ctx, star)
    t, v, s, pp, _ = g.greedy_pp(4)
    assert (t, v, len(s), pp) == (0, 0, 0, [(0, 0)] * 4)
    # a graph with no vertices (an empty induced subgraph)
    e, _ = g.induced(np.zeros(g.n, np.uint8))
    t, v, s, pp, _ = e.greedy_pp(2)
    assert (t, v, len(s), pp) == (0, 0, 0, [(0, 0)] * 2)
    e.close()
    g.close()


@pytest.mark.parametrize("n", [31, 32, 33, 1023, 1024, 1025, 1120, 1121, 2049])
def test_greedy_random_word_boundaries(ctx, n):
    """n around multiples of 32 (bitmap words) and 1,024 (lanes per workgroup), and at the largest graph whose adjacency
    bitmap fits 160 KB of LDS (1,120 vertices: the dense form) and one more (the sparse form); the last id is in an edge so
    n holds."""
    rng = np.random.default_rng(900 + n)
    p = min(0.5, 12.0 / n) if n > 64 else 0.4
    raw = np.concatenate([random_graph(rng, n, p, dup=0.2, loops=1), np.array([[0, n - 1], [1, n - 1], [0, 1]], np.uint32)])
    check(ctx, raw, (1, 2, 7), f"n={n}")


@pytest.mark.parametrize("seed", range(4))
def test_greedy_random(ctx, seed):
    rng = np.random.default_rng(4200 + seed)
    n = int(rng.integers(50, 3000))
    raw = random_graph(rng, n, float(rng.uniform(2.0, 40.0) / n), dup=0.3, loops=3)
    check(ctx, raw, (1, 3, 10), seed)


@pytest.mark.parametrize("n,p", [(700, 0.4), (1100, 0.15)])
def test_greedy_dense_random(ctx, n, p):
    """Dense graphs like the pruned ones (hundreds of neighbours per vertex, millions of triangles)."""
    rng = np.random.default_rng(n)
    check(ctx, random_graph(rng, n, p), (1, 5), (n, p))


def test_greedy_powerlaw(ctx):
    check(ctx, gen.powerlaw_small(), (1, 10), "powerlaw_small")


def test_greedy_on_pruned_graphs(ctx):
    """The pipeline of gpu_dsg.cu:318-356: core numbers, prune at ceil(kmax-core density), Greedy++ on what is left."""
    rng = np.random.default_rng(77)
    for trial in range(3):
        raw = np.concatenate([gen.powerlaw_small(), random_graph(rng, 5000, 0.002)])
        if trial:
            base = 9000 + 40 * trial
            raw = np.concatenate([raw, np.array(clique(30 + 10 * trial, base), np.uint32)])
        g = ga.DeviceGraph.from_pairs(ctx, raw)
        c = ga.Core(g, True)
        core, _, kmax = c.run()
        T_k, _ = g.vertex_participation((core >= kmax).astype(np.uint8))
        verts = int((core >= kmax).sum())
        k = (T_k + verts - 1) // verts
        sub, _, _ = c.prune(k)
        og = oracle.OGraph.from_pairs(raw)
        osub, _, _ = og.prune_to_core(core, k)
        for p in (1, 10):
            same(sub.greedy_pp(p), osub.greedy_pp(p), (trial, p))
        sub.close()
        c.close()
        g.close()


def test_greedy_size_limit_and_arguments(ctx):
    cap = ctx.greedy_max_vertices()
    assert cap >= 4096
    rng = np.random.default_rng(5)
    raw = np.concatenate([random_graph(rng, cap, 6.0 / cap), np.array([[0, cap - 1], [1, cap - 1], [0, 1]], np.uint32)])
    check(ctx, raw, (2,), "n = limit")
    big = np.concatenate([raw, np.array([[2, cap]], np.uint32)])   # one vertex more: the HBM form
    check(ctx, big, (1, 2), "n = limit + 1")
    g = ga.DeviceGraph.from_pairs(ctx, np.array(clique(5), np.uint32))
    with pytest.raises(ga.GAError) as ei:
        g.greedy_pp(0)
    assert ei.value.code == ga.GA_ERR_ARG
    g.close()


def test_greedy_many_equals_one_by_one(ctx):
    """ga_greedy_pp_many: every graph's best density and set equal ga_greedy_pp's and the oracle's, for a mix of dense
    graphs (one launch), graphs above the LDS-bitmap size (their own launches), triangle-free and empty graphs; twice in a
    row with a count in between."""
    rng = np.random.default_rng(31)
    raws = []
    for i in range(40):
        n = int(rng.integers(5, 900))
        raws.append(random_graph(rng, n, float(rng.uniform(0.02, 0.5)), dup=0.1, loops=1))
    raws[3] = np.array(clique(12), np.uint32)
    raws[5] = np.array([(0, i) for i in range(1, 40)], np.uint32)            # no triangle
    raws[7] = np.zeros((0, 2), np.uint32)                                    # no edge: n = 1 after the build
    raws[9] = np.concatenate([random_graph(rng, 1500, 0.01), np.array([[0, 1499]], np.uint32)])   # CSR form
    # dense form with a second owned vertex per lane (1,024 < n <= 1,120): its triangles come from the LDS bitmap
    raws[10] = np.concatenate([random_graph(rng, 1100, 0.1), np.array([[0, 1099]], np.uint32)])
    raws[15] = np.concatenate([random_graph(rng, 1120, 0.05), np.array([[3, 1119]], np.uint32)])
    graphs = [ga.DeviceGraph.from_pairs(ctx, r) for r in raws]
    assert graphs[10].n == 1100 and graphs[15].n == 1120
    for rep in range(2):
        st, sets, ms = ctx.greedy_pp_many(graphs, 4)
        for i, (g, r) in enumerate(zip(graphs, raws)):
            t, v, s, _, _ = g.greedy_pp(4)
            assert (int(st[i][0]), int(st[i][1])) == (t, v) and np.array_equal(sets[i], s), (rep, i)
            if i % 5 == 0 or i in (10, 15):
                ot, ov, os_, _ = oracle.OGraph.from_pairs(r).greedy_pp(4)
                assert (t, v) == (ot, ov) and np.array_equal(s, os_), (rep, i)
        graphs[0].count()
    for g in graphs:
        g.close()


def test_greedy_large_graphs(ctx):
    """Graphs above the LDS limit (the whole graph of gpu_dsg.cu:318-330), where the state lives in HBM and the minimum
    comes from block minima: random graphs, a power-law graph with hubs and many vertices in no triangle (the key-0
    bulk at the start of each pass), planted cliques (ties of equal keys across blocks and super-blocks), and ids
    spread past one super-block of 4,096."""
    cap = ctx.greedy_max_vertices()
    rng = np.random.default_rng(41)
    cases = [
        ("random", np.concatenate([random_graph(rng, cap + 3000, 8.0 / cap), np.array([[0, cap + 2999]], np.uint32)])),
        ("chung_lu", gen.chung_lu(30000, 150000, 2.1, 0.0, seed=9)),
        ("cliques", np.concatenate([np.array(clique(20, 4090), np.uint32), np.array(clique(20, 9000), np.uint32),
                                    np.array(clique(12, 20000), np.uint32), random_graph(rng, 24000, 0.0002)])),
    ]
    for tag, raw in cases:
        g = ga.DeviceGraph.from_pairs(ctx, raw)
        assert g.n > cap, tag
        og = oracle.OGraph.from_pairs(raw)
        for p in (1, 3):
            got = g.greedy_pp(p)
            same(got, og.greedy_pp(p), (tag, p))
        print(f"greedy_pp large {tag}: n {g.n}, m {g.m}, 3 passes in {got[4]:.1f} ms", flush=True)
        g.close()
