"""Pins the CPU oracle (oracle/oracle.hpp) before it is trusted as the checker of the GPU path:
golden vectors written by the reference's own code, the reference build itself (oracle/_ref) on fresh inputs,
closed forms, brute force, and invariants chosen so that a dropped term, a wrong index or a flipped rule fails."""
import json
import math
import os

import numpy as np
import pytest

import oracle
from conftest import DIVERGENT_TEXT_CASES, TEXT_CASES, clique, random_graph, zoo

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def test_golden_vectors(golden):
    """Every oracle output equals the reference's on the committed fixtures (loader, CSR, counts, peel, core)."""
    assert len(golden["graphs"]) >= 30
    for rec in golden["graphs"]:
        raw = np.array(rec["raw_pairs"], dtype=np.uint32).reshape(-1, 2)
        g = oracle.OGraph.from_pairs(raw)
        r, c, s = g.arrays()
        assert g.n == rec["n"] and g.m == rec["m"] and g.max_out_degree == rec["max_out_degree"], rec["name"]
        assert r.tolist() == rec["rowptr"] and c.tolist() == rec["col"] and s.tolist() == rec["src"], rec["name"]
        assert g.triangles() == rec["T"] and g.triangles_omp() == rec["T"]
        T, tri = g.participation()
        assert T == rec["T"] and tri.tolist() == rec["tri"]
        T2, tri2 = g.participation(omp=True)
        assert T2 == rec["T"] and tri2.tolist() == rec["tri"]
        Ts, sup = g.edge_support()
        assert Ts == rec["T"] and sup.tolist() == rec["sup"]
        core, kmax = g.core()
        assert core.tolist() == rec["core"] and kmax == rec["kmax"]
        assert list(g.core_level_density(core, kmax)) == rec["kmax_core"]
        assert g.edge_order().tolist() == rec["edge_order"]
        for key, p in rec["peel"].items():
            num, den = map(int, key.split("/"))
            bn, bd, br, rounds, rr = g.batch_peel(num, den)
            assert (bn, bd, br, rounds) == (p["best_num"], p["best_den"], p["best_round"], p["rounds"]), (rec["name"], key)
            assert rr.tolist() == p["removed_round"]


def test_text_loader_matches_reference(tmp_path):
    """The oracle's text loader equals the reference's on comments, blank lines, one-number lines, loops, duplicates."""
    if not oracle.ref_available():
        pytest.skip("reference build absent")
    path = tmp_path / "g.txt"
    path.write_text("# c\n% c\n\n1 2\n2 1\n3 3\n4\t5\n5 4 extra\n7\n 8 9\n10 x\n9 8\n")
    o = oracle.OGraph.load_text(str(path))
    r = oracle.RGraph.load_text(str(path))
    assert (o.n, o.m) == (r.n, r.m) == (10, 3)
    for a, b in zip(o.arrays(), r.arrays()):
        assert np.array_equal(a, b)


def test_text_loader_corner_cases_match_reference(tmp_path):
    """The oracle's loader equals the reference's own load_graph on every text corner case the GPU ingest is tested on."""
    if not oracle.ref_available():
        pytest.skip("reference build absent")
    for name, text in TEXT_CASES.items():
        path = tmp_path / "g.txt"
        path.write_bytes(text)
        o = oracle.OGraph.load_text(str(path))
        r = oracle.RGraph.load_text(str(path))
        assert (o.n, o.m) == (r.n, r.m), name
        for a, b in zip(o.arrays(), r.arrays()):
            assert np.array_equal(a, b), name


def test_loader_divergence_documented(tmp_path):
    """The one loader input class where this library departs from the reference on purpose: a pair that repeats
    another only after the uint32 cast is kept once here and twice by the reference."""
    for name, (text, n, m) in DIVERGENT_TEXT_CASES.items():
        path = tmp_path / "g.txt"
        path.write_bytes(text)
        o = oracle.OGraph.load_text(str(path))
        assert (o.n, o.m) == (n, m), name
        if oracle.ref_available():
            r = oracle.RGraph.load_text(str(path))
            assert r.n == n and r.m == m + 1, name   # the reference's repeated edge


@pytest.mark.parametrize("seed", range(6))
def test_oracle_equals_reference_build(seed):
    """Fresh seeded graphs (not in the fixtures): every oracle function equals the reference build."""
    if not oracle.ref_available():
        pytest.skip("reference build absent")
    rng = np.random.default_rng(1000 + seed)
    n = int(rng.integers(50, 700))
    raw = random_graph(rng, n, float(rng.uniform(0.01, 0.12)), dup=0.3, loops=2)
    o = oracle.OGraph.from_pairs(raw)
    uniq = oracle.canonical_unique(raw)
    r = oracle.RGraph.build(o.n, uniq)
    for a, b in zip(o.arrays(), r.arrays()):
        assert np.array_equal(a, b)
    assert o.triangles() == r.triangles() == r.triangles_omp()
    for p in (1.0, 0.9, 0.6, 0.3, 0.05, 0.0):
        alive = (rng.random(o.n) < p).astype(np.uint8)
        To, to = o.participation(alive)
        Tr, tr = r.participation(alive)
        assert To == Tr and np.array_equal(to, tr)
    assert o.edge_support()[1].tolist() == r.edge_support()[1].tolist()
    for num, den in ((1, 10), (1, 100), (1, 20), (3, 7)):
        assert o.batch_peel(num, den)[:4] == r.batch_peel(num, den)[:4]
        assert np.array_equal(o.batch_peel(num, den)[4], r.batch_peel(num, den)[4])
    co, ko = o.core()
    cr, kr = r.core()
    assert np.array_equal(co, cr) and ko == kr
    assert np.array_equal(o.edge_order(), r.edge_order())
    for a, b in zip(o.undirected(), r.undirected()):
        assert np.array_equal(a, b)
    for chunk in (1, 3, 128):   # build_tasks (graph.h:138-150)
        for a, b in zip(o.build_tasks(chunk), r.build_tasks(chunk)):
            assert np.array_equal(a, b), chunk


def test_golden_prune(golden):
    """prune_to_core of the oracle equals the reference's on every fixture graph at k in {0, 1, ceil(kmax/2), kmax, kmax+1}."""
    cases = 0
    for rec in golden["graphs"]:
        raw = np.array(rec["raw_pairs"], dtype=np.uint32).reshape(-1, 2)
        g = oracle.OGraph.from_pairs(raw)
        core = np.array(rec["core"], dtype=np.uint32)
        for p in rec["prune"]:
            pg, old, cout = g.prune_to_core(core, p["k"])
            r, c, _ = pg.arrays()
            assert (pg.n, pg.m, pg.max_out_degree) == (p["n"], p["m"], p["max_out_degree"]), (rec["name"], p["k"])
            assert r.tolist() == p["rowptr"] and c.tolist() == p["col"], (rec["name"], p["k"])
            assert old.tolist() == p["old_id"] and cout.tolist() == p["core_out"], (rec["name"], p["k"])
            cases += 1
    assert cases >= 120


@pytest.mark.parametrize("seed", range(4))
def test_induced_and_prune_equal_reference_build(seed):
    """Fresh seeded graphs and random keep masks: induced_subgraph and prune_to_core equal the reference build."""
    if not oracle.ref_available():
        pytest.skip("reference build absent")
    rng = np.random.default_rng(2000 + seed)
    n = int(rng.integers(80, 600))
    raw = random_graph(rng, n, float(rng.uniform(0.02, 0.15)), dup=0.2, loops=1)
    o = oracle.OGraph.from_pairs(raw)
    r = oracle.RGraph.build(o.n, oracle.canonical_unique(raw))
    for p in (1.0, 0.7, 0.3, 0.0):
        keep = (rng.random(o.n) < p).astype(np.uint8)
        (og, oo), (rg, ro) = o.induced(keep), r.induced(keep)
        assert np.array_equal(oo, ro) and (og.n, og.m) == (rg.n, rg.m)
        for a, b in zip(og.arrays(), rg.arrays()):
            assert np.array_equal(a, b)
    core, kmax = o.core()
    for k in (0, 1, kmax // 3, kmax, kmax + 1):
        op, oo, oc = o.prune_to_core(core, k)
        rp, ro, rc = r.prune_to_core(core, k)
        assert np.array_equal(oo, ro) and np.array_equal(oc, rc)
        for a, b in zip(op.arrays(), rp.arrays()):
            assert np.array_equal(a, b)


def test_prune_invariants():
    """What the mathematics fixes: keeping every vertex rebuilds the same CSR; the induced subgraph's T is the alive-masked
    count of the full graph (count_triangles_in, dsg.h:90-94); the pruned graph's own core numbers equal the restricted
    ones (the containment argument of dsg.h:629-635, checked by gpu_dsg.cu:325-328); new ids ascend with old ids."""
    rng = np.random.default_rng(77)
    for trial in range(6):
        n = int(rng.integers(100, 500))
        o = oracle.OGraph.from_pairs(random_graph(rng, n, float(rng.uniform(0.03, 0.12))))
        full, old = o.induced(np.ones(o.n, np.uint8))
        assert np.array_equal(old, np.arange(o.n))
        for a, b in zip(full.arrays(), o.arrays()):
            assert np.array_equal(a, b)
        keep = (rng.random(o.n) < 0.6).astype(np.uint8)
        sub, old = o.induced(keep)
        assert np.array_equal(old, np.flatnonzero(keep)) and np.all(np.diff(old.astype(np.int64)) > 0)
        assert sub.triangles() == o.participation(keep)[0]
        core, kmax = o.core()
        for k in (1, max(kmax // 2, 1), kmax):
            pg, old, cout = o.prune_to_core(core, k)
            assert np.array_equal(old, np.flatnonzero(core >= k))
            assert np.all(cout >= k)
            own, _ = pg.core()
            assert np.array_equal(own, cout), (trial, k)
    # the empty selection is a graph with no vertices
    e, old = o.induced(np.zeros(o.n, np.uint8))
    assert (e.n, e.m, old.size) == (0, 0, 0)


def _truss_by_pair(g):
    t, kmax = g.truss()
    r, c, s = g.arrays()
    return {(int(min(a, b)), int(max(a, b))): int(x) for a, b, x in zip(s, c, t)}, kmax


@pytest.mark.parametrize("seed", range(10))
def test_truss_brute_force(seed):
    """k-truss peeling equals the definition (repeated removal of edges with too few triangles per level)."""
    rng = np.random.default_rng(900 + seed)
    n = int(rng.integers(6, 36))
    raw = random_graph(rng, n, float(rng.uniform(0.1, 0.7)))
    if seed % 3 == 0:   # plant a clique so that high levels occur
        k = int(rng.integers(4, min(n, 9)))
        raw = np.concatenate([raw, np.array(clique(k), np.uint32)])
    g = oracle.OGraph.from_pairs(raw)
    uniq = oracle.canonical_unique(raw)
    bt = oracle.brute_truss(g.n, uniq)
    want = {(int(a), int(b)): int(x) for (a, b), x in zip(uniq, bt)}
    got, kmax = _truss_by_pair(g)
    assert got == want
    assert kmax == (max(want.values()) if want else 0)


def test_truss_closed_forms_and_invariants():
    """K_k: every edge k; trees and cycles (no triangle): 2; wheels W_n (n >= 5): 3; two K_k sharing an edge: k for
    all, including the shared edge; truss <= support + 2; every edge of the j-truss closes >= j - 2 triangles in it."""
    for k in range(3, 12):
        t, kmax = oracle.OGraph.from_pairs(np.array(clique(k), np.uint32)).truss()
        assert kmax == k and np.all(t == k)
    tree = np.array([(i, (i - 1) // 2) for i in range(1, 40)], np.uint32)
    cyc = np.array([(i, (i + 1) % 30) for i in range(30)], np.uint32)
    for raw in (tree, cyc):
        t, kmax = oracle.OGraph.from_pairs(raw).truss()
        assert kmax == 2 and np.all(t == 2)
    for n in range(5, 12):
        wheel = np.array([e for i in range(1, n) for e in ((0, i), (i, i % (n - 1) + 1))], np.uint32)
        t, kmax = oracle.OGraph.from_pairs(wheel).truss()
        assert kmax == 3 and np.all(t == 3), n
    for k in (4, 6, 9):
        two = sorted(set(clique(k)) | set(clique(k, k - 2)))
        t, kmax = oracle.OGraph.from_pairs(np.array(two, np.uint32)).truss()
        assert kmax == k and np.all(t == k)
    rng = np.random.default_rng(5)
    for _ in range(4):
        g = oracle.OGraph.from_pairs(random_graph(rng, 300, 0.08))
        t, kmax = g.truss()
        _, sup = g.edge_support()
        assert np.all(t <= sup + 2) and np.all(t >= 2)
        r, c, s = g.arrays()
        for j in range(3, kmax + 1):
            keep = t >= j
            sub = oracle.OGraph.from_pairs(np.stack([s[keep], c[keep]], 1), g.n)
            _, ss = sub.edge_support()
            assert ss.size == 0 or ss.min() + 2 >= j, j


def test_closed_forms():
    for k in range(3, 16):
        g = oracle.OGraph.from_pairs(np.array(clique(k), np.uint32))
        T = g.triangles()
        assert T == math.comb(k, 3)
        _, sup = g.edge_support()
        assert set(sup.tolist()) == {k - 2}                       # every edge of K_k lies in k-2 triangles
        _, tri = g.participation()
        assert set(tri.tolist()) == {math.comb(k - 1, 2)}
        core, kmax = g.core()
        assert set(core.tolist()) == {math.comb(k - 1, 2)} and kmax == math.comb(k - 1, 2)
        bn, bd, br, rounds, rr = g.batch_peel(1, 10)
        assert (bn, bd, br, rounds) == (math.comb(k, 3), k, 0, 1) and set(rr.tolist()) == {0}   # K_k peels in round 0
    for a, b in ((3, 4), (5, 5), (2, 9)):
        g = oracle.OGraph.from_pairs(np.array([(i, a + j) for i in range(a) for j in range(b)], np.uint32))
        assert g.triangles() == 0                                  # bipartite: no triangle
    for c in range(4, 20):                                         # wheel: hub + c-cycle has c triangles
        e = [(0, i) for i in range(1, c + 1)] + [(i, i % c + 1) for i in range(1, c + 1)]
        assert oracle.OGraph.from_pairs(np.array(e, np.uint32)).triangles() == c
    # disjoint K_a + K_b (a > b): the larger clique survives one round longer and is the best set
    e = clique(6) + clique(4, 6)
    bn, bd, br, rounds, rr = oracle.OGraph.from_pairs(np.array(e, np.uint32)).batch_peel(1, 10)
    assert (bn, bd) == (math.comb(6, 3), 6) and rr[:6].max() > rr[6:].max()


@pytest.mark.parametrize("seed", range(8))
def test_brute_force_counts(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(20, 400))
    raw = random_graph(rng, n, float(rng.uniform(0.02, 0.3)), dup=0.2, loops=1)
    g = oracle.OGraph.from_pairs(raw)
    uniq = oracle.canonical_unique(raw)
    T = oracle.brute_triangles(g.n, uniq)
    assert g.triangles() == T == g.triangles_omp()
    _, sup = g.edge_support()
    assert int(sup.sum()) == 3 * T
    for p in (1.0, 0.7, 0.3, 0.0):
        alive = (rng.random(g.n) < p).astype(np.uint8)
        Tb, tb = oracle.brute_participation(g.n, uniq, alive)
        To, to = g.participation(alive)
        assert To == Tb and np.array_equal(to, tb) and int(to.sum()) == 3 * To


def test_brute_force_support():
    rng = np.random.default_rng(5)
    raw = random_graph(rng, 120, 0.15)
    g = oracle.OGraph.from_pairs(raw)
    r, c, s = g.arrays()
    adj = np.zeros((g.n, g.n), bool)
    adj[s, c] = adj[c, s] = True
    _, sup = g.edge_support()
    for e in range(g.m):
        u, v = s[e], c[e]
        assert sup[e] == int(np.sum(adj[u] & adj[v]))


def test_tiny_exhaustive_peel_and_core():
    """n <= 12: core numbers by definition over all subsets, fixpoint definition and Alg 3 agree; the peel result is
    within 3(1+eps) of the brute-force optimum and never above it (the reference's tiny-suite checks)."""
    rng = np.random.default_rng(777)
    graphs = [(name, n, e) for name, n, e in zoo() if n <= 12]
    for it in range(120):
        n = int(rng.integers(3, 13))
        graphs.append((f"rand{it}", n, oracle.canonical_unique(random_graph(rng, n, float(rng.uniform(0.1, 0.9))))))
    with_tri = 0
    for name, n, e in graphs:
        g = oracle.OGraph.from_pairs(e, n)
        core, _ = g.core()
        assert np.array_equal(core, oracle.brute_core(n, e)), name
        assert np.array_equal(core, g.core_fixpoint()), name
        c2, rounds, _ = g.core_levels()
        assert np.array_equal(core, c2) and rounds >= (1 if n else 0)
        on, od = oracle.brute_densest(n, e)
        with_tri += on > 0
        for num, den in ((1, 10), (1, 2), (1, 1), (3, 1), (0, 1)):
            bn, bd, br, rounds, rr = g.batch_peel(num, den)
            assert bn * od <= on * bd                             # never above the optimum
            assert 3 * (den + num) * bn * od >= on * bd * den      # 3(1+eps) guarantee
            assert (rr < rounds).all()
    assert with_tri > 60


def test_arith_guards():
    assert oracle.peel_arith_ok(10, 5, 1, 10)
    assert not oracle.peel_arith_ok(10, 5, 1, 0)
    assert not oracle.peel_arith_ok(10, 5, 1, (1 << 20) + 1)
    assert not oracle.peel_arith_ok(10, 5, (1 << 20) + 1, 10)
    assert not oracle.peel_arith_ok(10, 1 << 32, 1, 10)
    assert not oracle.peel_arith_ok(4_000_000_000, 4_000_000_000, 1, 1 << 20)
    if oracle.ref_available():
        L = oracle.rlib()
        for args in ((10, 5, 1, 10), (10, 5, 1, 0), (10, 1 << 32, 1, 10), (4_000_000_000, 4_000_000_000, 1, 1 << 20), (7, 3, 0, 1)):
            assert bool(L.ref_peel_arith_ok(*args)) == oracle.peel_arith_ok(*args)


def test_planted_mistake_is_caught():
    """A mutated restatement (orientation rule flipped to the larger (degree, id) end) must fail the checks above:
    rows stop being the reference's and the per-edge index sets change, although T is orientation-independent."""
    rng = np.random.default_rng(3)
    raw = random_graph(rng, 200, 0.08)
    g = oracle.OGraph.from_pairs(raw)
    r, c, s = g.arrays()
    # flipped orientation built from the same edges
    flipped_src, flipped_dst = c.copy(), s.copy()
    order = np.lexsort((flipped_dst, flipped_src))
    fr = np.zeros(g.n + 1, np.uint32)
    np.add.at(fr, flipped_src.astype(np.int64) + 1, 1)
    fr = np.cumsum(fr).astype(np.uint32)
    gf = oracle.OGraph.from_csr(fr, flipped_dst[order])
    assert gf.triangles() == g.triangles()                        # T is invariant ...
    assert not np.array_equal(gf.arrays()[0], r)                  # ... but the CSR is not the reference's


def test_golden_greedy(golden):
    """greedy_pp of the oracle equals the reference's with 1, 3 and 10 passes on every fixture graph: best density, its
    vertex set and the best after every pass."""
    cases = 0
    for rec in golden["graphs"]:
        g = oracle.OGraph.from_pairs(np.array(rec["raw_pairs"], dtype=np.uint32).reshape(-1, 2))
        for passes, exp in rec["greedy_pp"].items():
            t, v, s, pp = g.greedy_pp(int(passes))
            assert (t, v, s.tolist()) == (exp["tri"], exp["verts"], exp["set"]), (rec["name"], passes)
            assert [list(x) for x in pp] == exp["per_pass"], (rec["name"], passes)
            cases += 1
    assert cases >= 90


@pytest.mark.parametrize("seed", range(3))
def test_greedy_equal_reference_build(seed):
    """Fresh seeded graphs, 1, 2 and 7 passes: the oracle's greedy_pp equals the reference build's."""
    if not oracle.ref_available():
        pytest.skip("reference build absent")
    rng = np.random.default_rng(3100 + seed)
    for _ in range(8):
        n = int(rng.integers(5, 400))
        raw = random_graph(rng, n, float(rng.uniform(0.02, 0.4)), dup=0.2, loops=1)
        o = oracle.OGraph.from_pairs(raw)
        r = oracle.RGraph.build(o.n, oracle.canonical_unique(raw))
        for passes in (1, 2, 7):
            a, b = o.greedy_pp(passes), r.greedy_pp(passes)
            assert a[:2] == b[:2] and np.array_equal(a[2], b[2]) and a[3] == b[3], (seed, n, passes)


def test_greedy_equal_reference_build_powerlaw():
    """The 8,298-vertex power-law fixture graph (thousands of re-keyed vertices per pass): the oracle's greedy_pp equals
    the reference build's at 1 and 3 passes."""
    if not oracle.ref_available():
        pytest.skip("reference build absent")
    from graph_analyzer_amd import gen
    raw = gen.powerlaw_small()
    o = oracle.OGraph.from_pairs(raw)
    r = oracle.RGraph.build(o.n, oracle.canonical_unique(raw))
    for passes in (1, 3):
        a, b = o.greedy_pp(passes), r.greedy_pp(passes)
        assert a[:2] == b[:2] and np.array_equal(a[2], b[2]) and a[3] == b[3], passes


def _set_triangles(uniq, s):
    keep = np.zeros(int(uniq.max()) + 1 if len(uniq) else 1, bool)
    keep[s[s < len(keep)]] = True
    sub = uniq[keep[uniq[:, 0]] & keep[uniq[:, 1]]] if len(uniq) else uniq
    return oracle.brute_triangles(len(keep), sub)


def test_greedy_against_exhaustive_optimum():
    """n <= 14, exhaustive optimum by brute force: Greedy++ is never denser than it; one pass (Charikar's peeling) is at
    least a third of it, the approximation bound of peeling for triangle density; the best never falls from pass to
    pass and the last pass's best is the result; the returned set holds exactly `tri` triangles on `verts` vertices."""
    rng = np.random.default_rng(515)
    for _ in range(60):
        n = int(rng.integers(3, 15))
        uniq = oracle.canonical_unique(random_graph(rng, n, float(rng.uniform(0.2, 0.95))))
        if not len(uniq):
            continue
        o = oracle.OGraph.from_pairs(uniq)
        on, od = oracle.brute_densest(o.n, uniq)
        t1, v1, s1, _ = o.greedy_pp(1)
        t, v, s, pp = o.greedy_pp(30)
        assert t * max(od, 1) <= on * max(v, 1)
        assert 3 * t1 * max(od, 1) >= on * max(v1, 1)
        dens = [a / b if b else 0.0 for a, b in pp]
        assert all(x <= y for x, y in zip(dens, dens[1:])) and pp[-1] == (t, v)
        for tt, vv, ss in ((t1, v1, s1), (t, v, s)):
            assert len(ss) == vv and _set_triangles(uniq, ss) == tt
            assert np.all(np.diff(ss.astype(np.int64)) > 0)


def test_greedy_closed_forms():
    """K_k: the whole clique, C(k,3)/k; disjoint K5 and K4: the K5 (10/5 beats 4/4 and the union's 14/9); graphs
    without a triangle and the empty graph: 0/0 and no set."""
    for k in (3, 4, 7, 12):
        o = oracle.OGraph.from_pairs(np.array(clique(k), np.uint32))
        for passes in (1, 5):
            t, v, s, pp = o.greedy_pp(passes)
            assert (t, v, s.tolist()) == (math.comb(k, 3), k, list(range(k)))
            assert pp == [(math.comb(k, 3), k)] * passes
    o = oracle.OGraph.from_pairs(np.array(clique(4) + clique(5, 4), np.uint32))
    t, v, s, _ = o.greedy_pp(3)
    assert (t, v, s.tolist()) == (10, 5, [4, 5, 6, 7, 8])
    for pairs in ([(0, 1), (1, 2), (2, 3), (3, 0)], [(0, i) for i in range(1, 9)], []):
        o = oracle.OGraph.from_pairs(np.array(pairs, np.uint32).reshape(-1, 2))
        t, v, s, pp = o.greedy_pp(4)
        assert (t, v, len(s)) == (0, 0, 0) and pp == [(0, 0)] * 4


def _greedy_model(n, pairs, passes, ge=False, use_load=True):
    """A small Python model of Greedy++ whose two rules can be broken on purpose: `ge` replaces the strict improvement
    test by >= (ties would keep the later, smaller subgraph), `use_load=False` drops the load term (every pass repeats
    Charikar's Greedy)."""
    adj = [set() for _ in range(n)]
    for a, b in pairs:
        adj[a].add(b)
        adj[b].add(a)
    tri0 = [sum(1 for i, a in enumerate(sorted(adj[v])) for b in sorted(adj[v])[i + 1:] if b in adj[a]) for v in range(n)]
    T0 = sum(tri0) // 3
    load = [0] * n
    bn = bd = 0
    best, per = None, []
    for _ in range(passes):
        cur, alive, order, T, improved, cut = tri0[:], [True] * n, [], T0, False, 0
        for left in range(n, 0, -1):
            a_, b_ = (bn, bd) if bd else (0, 1)
            if ((T * b_ >= a_ * left) if ge else (T * b_ > a_ * left)) and not (ge and T == 0):
                bn, bd, improved, cut = T, left, True, len(order)
            v = min((x for x in range(n) if alive[x]), key=lambda x: ((load[x] if use_load else 0) + cur[x], x))
            nb = [a for a in adj[v] if alive[a]]
            for i, a in enumerate(nb):
                for b in nb[i + 1:]:
                    if b in adj[a]:
                        cur[a] -= 1
                        cur[b] -= 1
            load[v] += cur[v]
            alive[v] = False
            T -= cur[v]
            order.append(v)
        if improved:
            best = (order[:], cut)
        per.append([bn, bd])
    return bn, bd, (sorted(best[0][best[1]:]) if bd else []), per


def test_greedy_fixtures_catch_broken_rules(golden):
    """The golden Greedy++ vectors are sensitive: a faithful model reproduces them on every fixture graph of at most 300
    vertices, while the same model with the tie rule flipped, or with the load term dropped, disagrees on at least one.
    So the golden comparison of the oracle would catch either mistake."""
    ge_differs = load_differs = 0
    for rec in golden["graphs"]:
        if rec["n"] > 300:
            continue
        g = oracle.OGraph.from_pairs(np.array(rec["raw_pairs"], dtype=np.uint32).reshape(-1, 2))
        _, c, s = g.arrays()
        pairs = list(zip(s.tolist(), c.tolist()))
        e = rec["greedy_pp"]["3"]
        want = (e["tri"], e["verts"], e["set"], e["per_pass"])
        assert _greedy_model(g.n, pairs, 3) == want, rec["name"]
        ge_differs += _greedy_model(g.n, pairs, 3, ge=True)[:3] != want[:3]
        load_differs += _greedy_model(g.n, pairs, 3, use_load=False) != want
    assert ge_differs >= 1 and load_differs >= 1



def test_exact_fixtures_pinned_by_brute_force():
    """The CoreExact / Exact fixtures (the reference's own results) on graphs of at most 20 vertices: every recorded best
    density equals the exhaustive maximum over all vertex subsets, and every recorded set holds the recorded number of
    triangles. On the larger fixtures the three searches (with and without Pruning1/2, seeded) agree with each other and
    never fall below the Greedy++ seed."""
    path = os.path.join(os.path.dirname(__file__), "golden", "exact_vectors.json")
    with open(path) as f:
        recs = json.load(f)["graphs"]
    checked = 0
    for rec in recs:
        raw = np.array(rec["raw_pairs"], np.uint32).reshape(-1, 2)
        og = oracle.OGraph.from_pairs(raw)
        results = [rec[k] for k in ("core_exact", "core_exact_no_pruning12", "core_exact_seeded", "exact_densest") if k in rec]
        for res in results:
            keep = np.zeros(og.n, np.uint8)
            keep[res["set"]] = 1
            assert og.participation(keep)[0] == res["tri"] and len(res["set"]) == res["verts"], rec["name"]
            assert res["tri"] * results[0]["verts"] == results[0]["tri"] * res["verts"], rec["name"]
        gp = rec["greedy_pp10"]
        assert results[0]["tri"] * max(gp["verts"], 1) >= gp["tri"] * results[0]["verts"], rec["name"]
        if og.n <= 20 and og.m:
            r, c, s = og.arrays()
            bn, bd = oracle.brute_densest(og.n, np.stack([s, c], 1))
            assert results[0]["tri"] * bd == bn * results[0]["verts"], rec["name"]
            checked += 1
    assert checked >= 25
