"""GPU parity: every device result equals the CPU oracle (itself pinned by tests/test_oracle.py) bit for bit.

Covers the GPU CSR builder (byte-identical rowptr/col/src), all count variants, per-vertex participation on random
alive masks, per-edge support, batch peeling (fused and host-driven) and the (k,triangle)-core (fused incremental,
host-driven incremental and recount), on the reference's fixtures, random graphs with duplicates and self loops,
hub graphs, empty and degenerate inputs, and several lane-group / LDS / block / chunk shapes."""
import json
import os

import numpy as np
import pytest

import graph_analyzer_amd as ga
import oracle
from conftest import clique, random_graph

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")
ALL_VARIANTS = [ga.TC_MERGE, ga.TC_GROUP4, ga.TC_GROUP8, ga.TC_GROUP16, ga.TC_GROUP64, ga.TC_HASH, ga.TC_SORTED, ga.TC_NOSRC,
                ga.TC_VERTEX]
EPS = [(1, 10), (1, 2), (1, 1), (3, 1), (0, 1)]


def check_graph(ctx, raw, n=None, eps=EPS, variants=ALL_VARIANTS, core=True, masks=True, tag=""):
    og = oracle.OGraph.from_pairs(raw, n)
    g = ga.DeviceGraph.from_pairs(ctx, raw, n)
    r, c, s = g.download()
    orr, oc, os_ = og.arrays()
    assert (g.n, g.m, g.max_out_degree) == (og.n, og.m, og.max_out_degree), tag
    assert np.array_equal(r, orr) and np.array_equal(c, oc) and np.array_equal(s, os_), f"CSR differs {tag}"
    T = og.triangles()
    for v in variants:
        assert g.count(v)[0] == T, f"variant {ga.variant_name(v)} {tag}"
    rng = np.random.default_rng(g.n + g.m)
    probs = (1.0, 0.9, 0.6, 0.3, 0.05, 0.0) if masks else (1.0,)
    for p in probs:
        alive = (rng.random(g.n) < p).astype(np.uint8)
        To, to = og.participation(alive)
        for v in (ga.TC_HASH, ga.TC_MERGE):
            Tg, tg = g.vertex_participation(alive, v)
            assert Tg == To and np.array_equal(tg, to), f"participation p={p} {ga.variant_name(v)} {tag}"
    Ts, so = og.edge_support()
    for v in (ga.TC_HASH, ga.TC_MERGE):
        Tg, sg = g.edge_support(v)
        assert Tg == Ts and np.array_equal(sg, so), f"edge support {ga.variant_name(v)} {tag}"
    for num, den in eps:
        bn, bd, br, rounds, rr = og.batch_peel(num, den)
        for fused in (True, False):
            pr = g.batch_peel(num, den, T_full=T, fused=fused)
            assert (pr.best_num, pr.best_den, pr.best_round, pr.rounds) == (bn, bd, br, rounds), f"peel {num}/{den} fused={fused} {tag}"
            assert np.array_equal(pr.removed_round, rr), f"removed_round {num}/{den} {tag}"
    if core:
        oc_, okmax = og.core()
        _, orounds, _ = og.core_levels()
        for inc, fused in ((True, True), (True, False), (False, False)):
            cg, rounds, kmax = g.core_decomposition(incremental=inc, fused=fused, T_full=T)
            assert np.array_equal(cg, oc_) and kmax == okmax, f"core inc={inc} fused={fused} {tag}"
            assert rounds == orounds, f"core rounds {rounds} vs {orounds} {tag}"
    g.close()
    return T


def test_golden_fixtures(ctx):
    """The reference's own outputs (tests/golden) reproduced on the GPU."""
    with open(GOLDEN) as f:
        recs = json.load(f)["graphs"]
    for rec in recs:
        raw = np.array(rec["raw_pairs"], dtype=np.uint32).reshape(-1, 2)
        g = ga.DeviceGraph.from_pairs(ctx, raw)
        r, c, s = g.download()
        assert r.tolist() == rec["rowptr"] and c.tolist() == rec["col"] and s.tolist() == rec["src"], rec["name"]
        for v in ALL_VARIANTS:
            assert g.count(v)[0] == rec["T"], (rec["name"], v)
        assert g.vertex_participation()[1].tolist() == rec["tri"]
        assert g.edge_support()[1].tolist() == rec["sup"]
        assert g.edge_order().tolist() == rec["edge_order"]
        og = oracle.OGraph.from_pairs(raw)
        ur, uc = g.undirected()   # build_undirected_csr (densest.h:162-171), against the oracle's
        our, ouc = og.undirected()
        assert np.array_equal(ur, our) and np.array_equal(uc, ouc), rec["name"]
        for chunk in (1, 5, 128):   # build_tasks (graph.h:138-150)
            for a, b in zip(g.build_tasks(chunk), og.build_tasks(chunk)):
                assert np.array_equal(a, b), (rec["name"], chunk)
        for key, p in rec["peel"].items():
            num, den = map(int, key.split("/"))
            pr = g.batch_peel(num, den, T_full=rec["T"])
            assert (pr.best_num, pr.best_den, pr.best_round, pr.rounds) == (p["best_num"], p["best_den"], p["best_round"], p["rounds"])
            assert pr.removed_round.tolist() == p["removed_round"]
        cg, rounds, kmax = g.core_decomposition(T_full=rec["T"])
        assert cg.tolist() == rec["core"] and kmax == rec["kmax"], rec["name"]
        g.close()


def test_closed_forms(ctx):
    import math
    for k in (3, 4, 7, 20, 70):
        e = np.array(clique(k), np.uint32)
        g = ga.DeviceGraph.from_pairs(ctx, e)
        for v in ALL_VARIANTS:
            assert g.count(v)[0] == math.comb(k, 3)
        assert set(g.edge_support()[1].tolist()) == {k - 2}
        core, rounds, kmax = g.core_decomposition()
        assert set(core.tolist()) == {math.comb(k - 1, 2)} and rounds == 1
        g.close()


def test_degenerate_inputs(ctx):
    # empty list -> n = 1, m = 0 (graph.h:76,94)
    for raw in (np.zeros((0, 2), np.uint32), np.array([[3, 3], [5, 5]], np.uint32)):
        g = ga.DeviceGraph.from_pairs(ctx, raw)
        assert (g.n, g.m) == (1, 0)
        assert g.count(ga.TC_HASH)[0] == 0 and g.count(ga.TC_MERGE)[0] == 0
        pr = g.batch_peel(1, 10, T_full=0)
        assert (pr.best_num, pr.best_den, pr.best_round, pr.rounds) == (0, 1, 0, 1)
        core, rounds, kmax = g.core_decomposition(T_full=0)
        assert core.tolist() == [0] and rounds == 1 and kmax == 0
        g.close()
    check_graph(ctx, np.array([[0, 1]], np.uint32), tag="single edge")
    check_graph(ctx, np.array([[5, 9]], np.uint32), tag="single edge, isolated low ids")
    # build_graph_from_edges semantics: n given, pairs must be unique and loop-free
    e = np.array([[0, 1], [1, 2], [0, 2]], np.uint32)
    g = ga.DeviceGraph.from_pairs(ctx, e, 10)
    assert (g.n, g.m, g.count()[0]) == (10, 3, 1)
    g.close()
    for bad in (np.array([[0, 1], [1, 0]], np.uint32), np.array([[0, 0]], np.uint32), np.array([[0, 12]], np.uint32)):
        with pytest.raises(ga.GAError) as ei:
            ga.DeviceGraph.from_pairs(ctx, bad, 10)
        assert ei.value.code == ga.GA_ERR_ARG


@pytest.mark.parametrize("seed", range(10))
def test_random_graphs(ctx, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(10, 1500))
    raw = random_graph(rng, n, float(rng.uniform(0.005, 0.2)), dup=0.3, loops=3)
    check_graph(ctx, raw, tag=f"seed {seed}")


@pytest.mark.parametrize("seed", range(4))
def test_powerlaw_graphs(ctx, seed):
    from graph_analyzer_amd import gen
    raw = gen.chung_lu(int(20000 + 10000 * seed), 150000, 2.0 + 0.15 * seed, 0.0, seed=seed + 11)
    check_graph(ctx, raw, eps=[(1, 10), (1, 100)], masks=True, tag=f"chung-lu {seed}")


def test_hub_graphs(ctx):
    """Hubs with 40-3000 neighbours: lane striding past 64 positions in the core update, hash items split by work."""
    rng = np.random.default_rng(99)
    for it in range(6):
        leaves = int(rng.integers(40, 3000))
        cl = int(rng.integers(8, 40))
        hubs = int(rng.integers(1, 4))
        n = hubs + leaves + cl
        perm = rng.permutation(n).astype(np.uint32)
        e = []
        c0 = hubs + leaves
        for a in range(c0, n):
            for b in range(a + 1, n):
                if rng.random() < 0.7:
                    e.append((a, b))
        for h in range(hubs):
            for l in range(leaves):
                if rng.random() < 0.9:
                    e.append((hubs + l, h))
            for _ in range(int(rng.integers(2, 6))):
                e.append((h, c0 + int(rng.integers(cl))))
            for l in range(0, leaves, 7):     # leaves that also close triangles through the hub
                e.append((hubs + l, hubs + (l + 1) % leaves))
        raw = perm[np.array(e, np.uint32)]
        check_graph(ctx, raw, eps=[(1, 10)], masks=False, tag=f"hub {it}")


def test_tuning_shapes(ctx):
    """Every lane-group width, block size, load factor, work chunk, flat chunk/block and item order gives the same exact
    count; the tuner too."""
    from graph_analyzer_amd import gen
    raw = gen.chung_lu(30000, 300000, 2.05, 0.0, seed=4)
    og = oracle.OGraph.from_pairs(raw)
    T = og.triangles()
    _, tri_ref = og.participation()
    _, sup_ref = og.edge_support()
    g = ga.DeviceGraph.from_pairs(ctx, raw)
    base = g.get_tuning()
    flats = ((0, 64), (2048, 256), (256, 64), (4096, 512), (4096, 1024))
    for li, lanes in enumerate(([4, 4, 8, 16, 32], [8, 16, 32, 64, 64], [16, 32, 64, 64, 64], [64, 64, 64, 64, 64], [4, 8, 16, 32, 64],
                                [1, 8, 32, 64, 64], [1, 16, 16, 32, 64])):   # 1: one lane per item for d+(h) <= 8
        for block in (64, 256, 1024):
            for chunk, shift, batches in ((64, 1, 1), (4096, 2, 8), (1 << 20, 3, 1024)):
                t = ga.TcTuning()
                t.lanes[:] = lanes
                t.block, t.work_chunk, t.load_shift, t.batches = block, chunk, shift, batches
                t.flat_chunk, t.flat_block = flats[(li + block) % len(flats)]
                t.item_order = (li + shift) % 3
                t.relabel = (li + block // 64 + shift) % 3   # the renumbered copy (2: line-aligned rows) for plain counts
                t.probe = (li + shift + block // 256) % 2   # lane slots or length classes for plain counts
                t.waves = (0, 5, 6, 7, 8, 4)[(li + block // 64 + shift) % 6]   # occupancy caps through LDS
                t.streams = (li + shift) % 2   # bucket launches overlapped on four streams
                for cap in (4096, 24, 3):   # small caps send large rows to the binary-search bucket
                    t.hash_cap = cap
                    g.set_tuning(t)
                    assert g.count(ga.TC_HASH)[0] == T, t.as_dict()
        for order in (0, 1, 2):
            t = g.get_tuning()
            t.item_order = order
            t.relabel = order   # participation and support stay on the CSR's own ids whatever relabel says
            g.set_tuning(t)
            assert np.array_equal(g.vertex_participation()[1], tri_ref), (lanes, order)
            assert np.array_equal(g.edge_support()[1], sup_ref), (lanes, order)
            assert g.count(ga.TC_HASH)[0] == T, (lanes, order)
    for fc, fb in flats + ((512, 128), (8192, 512), (1024, 64), (16384, 1024)):
        t = g.get_tuning()
        t.hash_cap, t.flat_chunk, t.flat_block = 4096, fc, fb
        try:
            g.set_tuning(t)
        except ga.GAError as e:   # a chunk whose LDS does not fit this device is refused up front
            assert e.code == ga.GA_ERR_ARG and fc >= 16384, (fc, fb)
            continue
        assert g.count(ga.TC_HASH)[0] == T, t.as_dict()
    g.set_tuning(base)
    best, ms = g.tune()
    assert g.count(ga.TC_HASH)[0] == T and ms > 0
    g.close()


# The multipliers of the lane-slot tables' rebuilds, in the order the kernel tries them.
SLOT_HASH_MULS = (2654435769, 2246822507, 3266489909, 668265263)


def colliding_ids(k, bbits=3):
    """k vertex ids that land in bucket 0 of a table of 2^bbits buckets under every multiplier (and so under fewer
    buckets too): a vertex with such out-neighbours leaves keys outside their home bucket in every table tried."""
    x = np.arange(1, 1 << 22, dtype=np.uint64)
    ok = np.ones(x.size, bool)
    for m in SLOT_HASH_MULS:
        ok &= (((x * np.uint64(m)) & np.uint64(0xFFFFFFFF)) >> np.uint64(32 - bbits)) == 0
    ids = x[ok][:k]
    assert ids.size == k
    return [int(v) for v in ids]


def test_slot_tables_that_never_settle(ctx):
    """Lane-slot counts read only a key's home bucket; a table in which a key left its home bucket is rebuilt with
    other multipliers, and a row whose every table still has such a key is counted by the per-entry walk. Rows whose
    out-neighbours collide under all multipliers (cliques of 9 and 6 on such ids, with the smallest id first so that it
    hashes the others) take that walk in the plain and the per-graph counts, at every load factor and occupancy cap."""
    ids = colliding_ids(13)
    verts = [0] + ids[:8]
    pairs = [(verts[a], verts[b]) for a in range(9) for b in range(a + 1, 9)]
    w = [7] + ids[8:13]
    pairs += [(w[a], w[b]) for a in range(6) for b in range(a + 1, 6)]
    raw = np.array(pairs, np.uint32)
    T = oracle.OGraph.from_pairs(raw).triangles()
    assert T == 84 + 20   # C(9,3) + C(6,3)
    g = ga.DeviceGraph.from_pairs(ctx, raw)
    other = random_graph(np.random.default_rng(5), 300, 0.08)
    T_other = oracle.OGraph.from_pairs(other).triangles()
    b = ga.Batch(ctx, [raw, other])   # the colliding graph first: its ids are unshifted in the union
    for lanes0 in (4, 8, 16):
        for shift in (1, 2, 3):
            for waves in (0, 6):
                t = ga.TcTuning.from_dict(ga.default_tuning().as_dict())
                t.lanes[0], t.load_shift, t.probe, t.relabel, t.waves = lanes0, shift, 1, 0, waves
                g.set_tuning(t)
                assert g.count(ga.TC_HASH)[0] == T, t.as_dict()
                b.set_tuning(t)
                assert [int(x) for x in b.count_all()[0]] == [T, T_other], t.as_dict()
    b.close()
    g.close()


def test_bucket_profile_work(ctx):
    """The per-launch profile accounts for every work entry of the plan once, and its probe elements add up to the sum
    of min(d+u, d+v) over the oriented edges (the B_alg term), whatever tuning splits them over the launches."""
    from graph_analyzer_amd import gen
    raw = gen.chung_lu(20000, 200000, 2.1, 1.0, seed=9)
    g = ga.DeviceGraph.from_pairs(ctx, raw)
    r, c, s = g.download()
    d = np.diff(r.astype(np.int64))
    mins = np.minimum(d[s], d[c])
    for flat, relabel, order in ((0, 0, 1), (2048, 0, 0), (0, 1, 2), (1024, 1, 1), (0, 2, 1), (2048, 2, 2)):
        t = g.get_tuning()
        t.flat_chunk, t.flat_block, t.relabel, t.item_order = flat, 256, relabel, order
        g.set_tuning(t)
        ent, el, ms = g.bucket_profile(2)
        nw, _ = g.plan_info()
        assert sum(ent) == nw == int((mins > 0).sum()), (flat, relabel)
        assert sum(el) == int(mins.sum()), (flat, relabel)
        assert all(x >= 0 for x in ms) and (ms[6] > 0) == (flat > 0), (flat, relabel, ms)
    g.close()


def test_relabel_counts_on_every_graph_kind(ctx):
    """The renumbered copy (relabel = 1, and its line-aligned rows, relabel = 2) with the windowed, XCD-contiguous item
    order counts every graph kind exactly, including graphs with isolated vertices, a star, cliques sharing vertices and
    triangle-free graphs, also when the tuning moves between the plain and the line-aligned rows (the plan's row starts
    are rewritten each way); the batch union keeps each graph's count."""
    from graph_analyzer_amd import gen
    rng = np.random.default_rng(77)
    cases = [gen.chung_lu(20000, 200000, 2.1, 2.0, seed=8), gen.communities(20000, 300, 10, 120, 0.3, seed=9),
             gen.rmat(14, 1 << 17, seed=10), np.array([(0, i) for i in range(1, 3000)], np.uint32),
             np.array(clique(40) + clique(40, 39) + clique(30, 5000), np.uint32), random_graph(rng, 700, 0.05, dup=0.2, loops=3)]
    for raw in cases:
        og = oracle.OGraph.from_pairs(raw)
        T = og.triangles()
        g = ga.DeviceGraph.from_pairs(ctx, raw)
        for relabel in (1, 2, 1, 2):
            for order in (1, 2):
                for lanes0 in (1, 4):
                    t = g.get_tuning()
                    t.relabel, t.item_order, t.lanes[0] = relabel, order, lanes0
                    g.set_tuning(t)
                    assert g.count(ga.TC_HASH)[0] == T, (relabel, order, lanes0)
        for relabel in (2, 1, 2):
            t = g.get_tuning()
            t.flat_chunk, t.flat_block, t.relabel = 2048, 256, relabel   # the flat kernel on the renumbered copy
            g.set_tuning(t)
            assert g.count(ga.TC_HASH)[0] == T, relabel
        t.hash_cap = 3   # rows past the cap: the binary-search bucket on the line-aligned rows
        g.set_tuning(t)
        assert g.count(ga.TC_HASH)[0] == T
        g.close()
    b = ga.Batch(ctx, cases)
    Tb, _ = b.count_all()
    for relabel in (1, 2):
        t = b.get_tuning()
        t.relabel, t.item_order = relabel, 2
        b.set_tuning(t)
        Tr, _ = b.count_all()
        Tp, _ = b.count_all(per_graph=True)
        assert np.array_equal(Tb, Tr) and np.array_equal(Tb, Tp), relabel
    assert [int(x) for x in Tb] == [oracle.OGraph.from_pairs(r).triangles() for r in cases]
    b.close()


def test_flat_kernel_edge_cases(ctx):
    """Light rows only, heavy rows only (K_200: d+(h) up to 199) and a mix; chunks smaller than one row's units."""
    from graph_analyzer_amd import gen
    for raw in (np.array(clique(12), np.uint32), np.array(clique(200), np.uint32), gen.powerlaw_small(),
                np.array([[0, 1], [1, 2], [0, 2]], np.uint32)):
        og = oracle.OGraph.from_pairs(raw)
        T = og.triangles()
        g = ga.DeviceGraph.from_pairs(ctx, raw)
        for fc, fb in ((256, 64), (512, 64), (2048, 256), (0, 64), (1024, 64)):
            t = g.get_tuning()
            t.flat_chunk, t.flat_block = fc, fb
            g.set_tuning(t)
            assert g.count(ga.TC_HASH)[0] == T, (g.n, fc, fb)
        g.close()


def test_peel_errors(ctx):
    e = np.array(clique(5), np.uint32)
    g = ga.DeviceGraph.from_pairs(ctx, e)
    for num, den in ((1, 0), (1, (1 << 20) + 1), ((1 << 20) + 1, 10)):
        with pytest.raises(ga.GAError) as ei:
            g.batch_peel(num, den, T_full=10)
        assert ei.value.code == ga.GA_ERR_ARITH
    with pytest.raises(ga.GAError) as ei:
        g.batch_peel(1, 10, T_full=11)           # wrong round-0 count is caught on the device
    assert ei.value.code == ga.GA_ERR_STATE
    with pytest.raises(ga.GAError):
        g.core_decomposition(T_full=11)
    g.close()


def test_powerlaw_small_workload(ctx):
    """The 8,298-vertex / 100,762-edge power-law workload end to end."""
    from graph_analyzer_amd import gen
    check_graph(ctx, gen.powerlaw_small(), eps=[(1, 10), (1, 100), (1, 20)], tag="powerlaw_small")


def test_core_many_small_graphs(ctx):
    """ga_core_run_many (one workgroup per graph, all rounds in one launch) equals the oracle and the fused loop on a
    mix of graphs: fixtures, cliques, random graphs with duplicates and loops, power-law graphs, degenerate ones."""
    from graph_analyzer_amd import gen
    rng = np.random.default_rng(5)
    raws = [np.array(clique(k), np.uint32) for k in (3, 9, 40)]
    raws += [np.zeros((0, 2), np.uint32), np.array([[2, 2]], np.uint32), np.array([[0, 1]], np.uint32)]
    raws += [random_graph(rng, int(rng.integers(20, 2500)), float(rng.uniform(0.002, 0.1)), dup=0.2, loops=2) for _ in range(10)]
    raws += [gen.chung_lu(int(rng.integers(2000, 30000)), int(rng.integers(20000, 200000)), 2.1, 0.0, seed=31 + i) for i in range(4)]
    with open(GOLDEN) as f:
        raws += [np.array(rec["raw_pairs"], dtype=np.uint32).reshape(-1, 2) for rec in json.load(f)["graphs"][:8]]
    graphs = [ga.DeviceGraph.from_pairs(ctx, r) for r in raws]
    Ts = [g.count(ga.TC_HASH)[0] for g in graphs]
    cores = [ga.Core(g, True) for g in graphs]
    for rep in range(2):   # twice on the same states: nothing may leak from one run to the next
        res, ms = ga.core_many(ctx, cores, Ts)
        for raw, g, (core, rounds, kmax) in zip(raws, graphs, res):
            og = oracle.OGraph.from_pairs(raw)
            oc, okmax = og.core()
            _, orounds, _ = og.core_levels()
            assert np.array_equal(core, oc) and kmax == okmax and rounds == orounds, (g.n, g.m, rep)
    fused = [g.core_decomposition(T_full=T) for g, T in zip(graphs, Ts)]
    for (core, rounds, kmax), (fc, fr, fk) in zip(res, fused):
        assert np.array_equal(core, fc) and (rounds, kmax) == (fr, fk)
    with pytest.raises(ga.GAError) as ei:   # a wrong given count is caught on the device
        ga.core_many(ctx, cores[:2], [Ts[0], Ts[1] + 1])
    assert ei.value.code == ga.GA_ERR_STATE
    for c in cores:
        c.close()
    for g in graphs:
        g.close()


def test_core_batch_shared_launches(ctx):
    """ga_core_run_batch (every round's launches shared by all unfinished states) equals the oracle and ga_core_run on
    graphs of very different sizes and round counts; twice in a row, the second time on a subset in another order, two
    states of one graph side by side, and with a wrong given count caught for the state it belongs to."""
    from graph_analyzer_amd import gen
    rng = np.random.default_rng(11)
    raws = [np.array(clique(k), np.uint32) for k in (4, 30)]
    raws += [np.zeros((0, 2), np.uint32), np.array([[0, 1]], np.uint32)]
    raws += [random_graph(rng, int(rng.integers(50, 3000)), float(rng.uniform(0.002, 0.05)), dup=0.1, loops=1) for _ in range(6)]
    raws += [gen.chung_lu(int(rng.integers(5000, 60000)), int(rng.integers(50000, 400000)), 2.1, 0.0, seed=71 + i) for i in range(3)]
    raws += [gen.rmat(14, 8 << 14, seed=5)]
    graphs = [ga.DeviceGraph.from_pairs(ctx, r) for r in raws]
    Ts = [g.count(ga.TC_HASH)[0] for g in graphs]
    want = []
    for raw in raws:
        og = oracle.OGraph.from_pairs(raw)
        oc, okmax = og.core()
        _, orounds, _ = og.core_levels()
        want.append((oc, orounds, okmax))
    cores = [ga.Core(g, True) for g in graphs] + [ga.Core(graphs[-1], True)]   # two states of the R-MAT graph
    idx = list(range(len(graphs))) + [len(graphs) - 1]
    res, ms = ga.core_batch(ctx, cores, [Ts[i] for i in idx])
    print(f"core batch: {len(cores)} states in {ms:.2f} ms; rounds {[r for _, r, _ in res]}", flush=True)
    for i, (core, rounds, kmax) in zip(idx, res):
        oc, orounds, okmax = want[i]
        assert np.array_equal(core, oc) and rounds == orounds and kmax == okmax, (i, graphs[i].n, graphs[i].m)
    sub = [len(cores) - 1, 10, 3, 0, 7]
    res2, _ = ga.core_batch(ctx, [cores[j] for j in sub], [Ts[idx[j]] for j in sub])
    for j, (core, rounds, kmax) in zip(sub, res2):
        oc, orounds, okmax = want[idx[j]]
        assert np.array_equal(core, oc) and rounds == orounds and kmax == okmax, ("second call", j)
    for j in (9, 10):   # equal to the single-graph fused loop
        fc, fr, fk = graphs[idx[j]].core_decomposition(T_full=Ts[idx[j]])
        assert np.array_equal(fc, res[j][0]) and (fr, fk) == res[j][1:]
    with pytest.raises(ga.GAError) as ei:   # a wrong given count is caught on the device, for that state only
        ga.core_batch(ctx, cores[8:11], [Ts[8], Ts[9] + 1, Ts[10]])
    assert ei.value.code == ga.GA_ERR_STATE and "core 1:" in str(ei.value)
    with pytest.raises(ga.GAError):
        ga.core_batch(ctx, [cores[0], cores[0]], [Ts[0], Ts[0]])
    for c in cores:
        c.close()
    for g in graphs:
        g.close()
