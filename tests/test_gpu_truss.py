"""GPU parity of the k-truss decomposition (ga_truss_*): every oriented edge's truss number and kmax equal the CPU
oracle's min-support-first peeling (itself pinned by brute force on the definition, closed forms and invariants in
tests/test_oracle.py), on the reference's fixtures, closed forms, random graphs with duplicates and self loops,
power-law graphs with planted cliques and degenerate inputs."""
import json
import os

import numpy as np
import pytest

import graph_analyzer_amd as ga
import oracle
from conftest import clique, random_graph

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")


def check(ctx, raw, tag):
    og = oracle.OGraph.from_pairs(raw)
    g = ga.DeviceGraph.from_pairs(ctx, raw)
    ot, okmax = og.truss()
    tr = ga.Truss(g)
    for rep in range(2):   # the same state twice: nothing may leak from one run into the next
        t, rounds, kmax = tr.run()
        assert kmax == okmax and np.array_equal(t, ot), (tag, rep)
        assert (rounds == 0) == (g.m == 0), tag
    tr.close()
    g.close()
    return okmax


def test_truss_fixtures(ctx):
    with open(GOLDEN) as f:
        for rec in json.load(f)["graphs"]:
            check(ctx, np.array(rec["raw_pairs"], dtype=np.uint32).reshape(-1, 2), rec["name"])


def test_truss_closed_forms(ctx):
    for k in (3, 4, 5, 8, 17, 40):
        assert check(ctx, np.array(clique(k), np.uint32), f"K{k}") == k
    for n in (5, 9, 30):
        wheel = np.array([e for i in range(1, n) for e in ((0, i), (i, i % (n - 1) + 1))], np.uint32)
        assert check(ctx, wheel, f"W{n}") == 3
    assert check(ctx, np.array([(i, (i - 1) // 2) for i in range(1, 300)], np.uint32), "tree") == 2
    assert check(ctx, np.zeros((0, 2), np.uint32), "no edges") == 0
    assert check(ctx, np.array([[3, 3]], np.uint32), "loop only") == 0
    assert check(ctx, np.array([[0, 1]], np.uint32), "one edge") == 2


@pytest.mark.parametrize("seed", range(6))
def test_truss_random(ctx, seed):
    rng = np.random.default_rng(700 + seed)
    n = int(rng.integers(30, 3000))
    raw = random_graph(rng, n, float(rng.uniform(0.003, 0.2)), dup=0.2, loops=2)
    if seed % 2 == 0:
        raw = np.concatenate([raw, np.array(clique(int(rng.integers(5, 30)), int(rng.integers(0, n // 2))), np.uint32)])
    check(ctx, raw, seed)


def test_truss_powerlaw(ctx):
    from graph_analyzer_amd import gen
    check(ctx, gen.powerlaw_small(), "powerlaw_small")
    check(ctx, gen.chung_lu(30000, 300000, 2.1, 0.0, seed=51), "chung_lu 30K")
    raw = np.concatenate([gen.chung_lu(80000, 600000, 2.3, 5.0, seed=52), gen.communities(80000, 400, 10, 60, 0.3, seed=53)])
    check(ctx, raw, "chung_lu 80K + communities")


def test_truss_rmat(ctx):
    """R-MAT scale 16 (477K edges, hubs of high degree: long shorter-lists split into chunks over waves)."""
    from graph_analyzer_amd import gen
    assert check(ctx, gen.rmat(16, 8 << 16, 0.57, 0.19, 0.19, 7), "rmat16") == 77
