"""Batching parity: every graph of a batch gets exactly the result it gets alone (oracle, per-graph launches, and a batch
of one), for the CSR, the count and the batch peel."""
import numpy as np
import pytest

import graph_analyzer_amd as ga
import oracle
from conftest import clique, random_graph

pytestmark = pytest.mark.gpu


def mix(count, seed, nmax=3000):
    from graph_analyzer_amd import gen
    rng = np.random.default_rng(seed)
    graphs = []
    for i in range(count):
        n = int(np.exp(rng.uniform(np.log(20), np.log(nmax))))
        deg = float(rng.uniform(2, 24))
        if i % 3 == 0:
            graphs.append(gen.erdos_renyi(n, int(n * deg / 2), seed=seed * 1000 + i))
        elif i % 3 == 1:
            graphs.append(gen.chung_lu(n, int(n * deg / 2), float(rng.uniform(2.0, 2.5)), 1.0, seed=seed * 1000 + i))
        else:
            graphs.append(random_graph(rng, n if n < 600 else 600, float(rng.uniform(0.01, 0.2)), dup=0.3, loops=2))
    graphs[1] = np.zeros((0, 2), np.uint32)                       # empty graph: n = 1
    graphs[2] = np.array([[4, 4], [7, 7]], np.uint32)             # loops only: n = 1
    graphs[3] = np.array(clique(30), np.uint32)
    return graphs


def test_batch_matches_oracle_and_per_graph(ctx):
    graphs = mix(48, 3)
    b = ga.Batch(ctx, graphs)
    assert b.count == len(graphs)
    T_batch, _ = b.count_all()
    T_each, _ = b.count_all(per_graph=True)
    assert np.array_equal(T_batch, T_each)
    oracles = [oracle.OGraph.from_pairs(p) for p in graphs]
    for i, og in enumerate(oracles):
        r, c, s = b.graph(i)
        orr, oc, os_ = og.arrays()
        assert np.array_equal(r, orr) and np.array_equal(c, oc) and np.array_equal(s, os_), i
        assert T_batch[i] == og.triangles(), i
    for num, den in ((1, 10), (1, 20)):
        stats, rr, _ = b.peel_all(num, den, T_full=T_batch)
        stats2, rr2, _ = b.peel_all(num, den, per_graph=True)
        assert np.array_equal(stats, stats2) and np.array_equal(rr, rr2)
        for i, og in enumerate(oracles):
            bn, bd, br, rounds, orr_ = og.batch_peel(num, den)
            assert tuple(int(x) for x in stats[i]) == (bn, bd, br, rounds), (i, num, den)
            assert np.array_equal(rr[b.voff[i]:b.voff[i + 1]], orr_), i
    b.close()


def test_alone_equals_in_mix(ctx):
    graphs = mix(24, 8)
    b = ga.Batch(ctx, graphs)
    T, _ = b.count_all()
    stats, rr, _ = b.peel_all(1, 10)
    for i in (0, 3, 5, 17, 23):
        one = ga.Batch(ctx, [graphs[i]])
        T1, _ = one.count_all()
        s1, rr1, _ = one.peel_all(1, 10)
        assert T1[0] == T[i] and np.array_equal(s1[0], stats[i]) and np.array_equal(rr1, rr[b.voff[i]:b.voff[i + 1]])
        g = ga.DeviceGraph.from_pairs(ctx, graphs[i])
        assert g.count()[0] == T[i]
        pr = g.batch_peel(1, 10)
        assert (pr.best_num, pr.best_den, pr.best_round, pr.rounds) == tuple(int(x) for x in stats[i])
        g.close()
        one.close()
    b.close()


def test_peel_rounds_captured_equal_direct(ctx):
    """The batched peel with its rounds captured into a HIP graph and replayed (1, 3, 8 and 64 rounds per graph) gives
    the same stats and removal rounds as every round enqueued directly, call after call, with eager counts between the
    calls; the direct form itself equals the oracle."""
    graphs = mix(32, 11)
    b = ga.Batch(ctx, graphs)
    T, _ = b.count_all()
    b.set_peel_graph(0)
    want = {eps: b.peel_all(*eps, T_full=T)[:2] for eps in ((1, 10), (1, 2), (0, 1))}
    for i in (0, 3, 7, 20):
        og = oracle.OGraph.from_pairs(graphs[i])
        bn, bd, br, rounds, orr_ = og.batch_peel(1, 10)
        assert tuple(int(x) for x in want[(1, 10)][0][i]) == (bn, bd, br, rounds), i
    for rounds in (1, 3, 8, 64):
        b.set_peel_graph(rounds)
        for eps, (st0, rr0) in want.items():
            for rep in range(2):
                st, rr, _ = b.peel_all(*eps, T_full=T)
                assert np.array_equal(st, st0) and np.array_equal(rr, rr0), (rounds, eps, rep)
                T2, _ = b.count_all()
                assert np.array_equal(T2, T)
    with pytest.raises(ga.GAError):
        b.set_peel_graph(65)
    b.close()


def test_batch_peel_errors(ctx):
    graphs = [np.array(clique(6), np.uint32), np.array(clique(5), np.uint32)]
    b = ga.Batch(ctx, graphs)
    with pytest.raises(ga.GAError) as ei:
        b.peel_all(1, 0)
    assert ei.value.code == ga.GA_ERR_ARITH
    with pytest.raises(ga.GAError) as ei:
        b.peel_all(1, 10, T_full=np.array([20, 11], np.uint64))   # graph 1 has 10 triangles
    assert ei.value.code == ga.GA_ERR_STATE and "graph 1" in str(ei.value)
    b.close()


def test_batch_tunings_flat_per_graph(ctx):
    """Per-graph counts of the flat kernel (chunks that straddle graph boundaries credit each hit to its own graph) and
    of every lane-group shape equal the oracle; the batch tuner times the per-graph launch and keeps results exact."""
    graphs = mix(64, 11, nmax=1500)
    ref = np.array([oracle.OGraph.from_pairs(p).triangles() for p in graphs], np.uint64)
    b = ga.Batch(ctx, graphs)
    for i, (lanes, fc, fb, order) in enumerate((([8, 16, 32, 64, 64], 256, 64, 1), ([4, 8, 32, 32, 64], 1024, 256, 0),
                                                ([4, 4, 16, 64, 64], 2048, 256, 1), ([8, 8, 64, 64, 64], 4096, 1024, 1),
                                                ([16, 32, 64, 64, 64], 0, 512, 0), ([1, 16, 32, 64, 64], 0, 256, 1),
                                                ([1, 8, 32, 64, 64], 2048, 256, 1), ([4, 16, 32, 64, 64], 0, 256, 1))):
        t = ga.TcTuning()
        t.lanes[:] = lanes
        t.block, t.work_chunk, t.batches, t.load_shift, t.hash_cap = 256, 2048, 4, 2, 4096
        t.flat_chunk, t.flat_block, t.item_order = fc, fb, order
        t.probe = i % 2   # per-graph counts in lane slots or by length class
        t.waves = (0, 4, 5, 6, 7, 8)[i % 6]   # occupancy caps through LDS
        b.set_tuning(t)
        T, _ = b.count_all()
        assert np.array_equal(T, ref), (lanes, fc, fb, order)
        Tp, _ = b.count_all(per_graph=True)
        assert np.array_equal(Tp, ref), (lanes, fc, fb, order)
    best, ms = b.tune()
    assert ms > 0 and np.array_equal(b.count_all()[0], ref)
    b.close()


def test_device_probes(ctx):
    """The measurement probes run, report what they did, and reject bad sizes."""
    ms = ctx.copy_probe(64 << 20)
    assert ms > 0
    cus = ctx.device_info()["cus"]
    for seg in (32, 64, 128, 256):
        n, ms = ctx.gather_probe(64 << 20, seg)
        assert ms > 0 and n == cus * 32 * 256 // (seg // 4) * 32 * 8, (seg, n, ms)
    with pytest.raises(ga.GAError):
        ctx.gather_probe(64 << 20, 96)
    for wg in (8, 64, cus):   # a device-wide barrier vs one launch boundary (the persistent-loop trade-off)
        bar_us, launch_us = ctx.barrier_probe(wg, 500)
        print(f"barrier probe: {wg} workgroups: {bar_us:.2f} us per barrier, {launch_us:.2f} us per launch", flush=True)
        assert bar_us > 0 and launch_us > 0
    with pytest.raises(ga.GAError):
        ctx.barrier_probe(cus * 1000, 10)   # cannot all be resident


def test_count_many_shared_launches(ctx):
    """ga_tc_count_many over separately resident graphs (one launch per bucket shape for all of them) equals each
    graph's own count and the oracle: graphs of every size and shape class, mixed tunings (one lane per item, lane
    groups in both probe forms, the renumbered copy, the binary-search bucket, and the flat kernel, which takes the
    per-graph path), empty and loop-only graphs, two calls in a row (the second on fewer graphs) and a capped grid."""
    graphs = mix(40, 21, nmax=4000)
    graphs.append(np.array(clique(60), np.uint32))
    ref = [oracle.OGraph.from_pairs(p).triangles() for p in graphs]
    dgs = [ga.DeviceGraph.from_pairs(ctx, p) for p in graphs]
    shapes = (([1, 16, 32, 64, 64], 0, 4096, 1, 0), ([4, 8, 16, 32, 64], 0, 4096, 0, 1), ([1, 32, 16, 64, 64], 0, 8, 2, 1),
              ([8, 16, 32, 64, 64], 1024, 4096, 1, 0))
    for i, g in enumerate(dgs):
        if i % 5 == 4:
            continue   # library defaults
        lanes, flat, cap, order, probe = shapes[i % 4]
        t = ga.TcTuning.from_dict(ga.default_tuning().as_dict())
        t.lanes[:] = lanes
        t.flat_chunk, t.flat_block = (flat, 256) if flat else (0, 256)
        t.hash_cap, t.item_order, t.probe, t.relabel = cap, order, probe, i % 3
        t.waves = (0, 5, 7, 6)[i % 4]
        t.streams = i % 2
        g.set_tuning(t)
    own = [g.count(ga.TC_HASH)[0] for g in dgs]
    assert own == ref
    for rep, sel in enumerate((list(range(len(dgs))), list(range(0, len(dgs), 3)))):
        T, ms = ctx.count_many([dgs[i] for i in sel], ga.TC_HASH)
        assert [int(x) for x in T] == [ref[i] for i in sel], rep
        assert ms > 0
    ctx.set_grid_cap(3)
    try:
        T, _ = ctx.count_many(dgs, ga.TC_HASH)
        assert [int(x) for x in T] == ref
    finally:
        ctx.set_grid_cap(0)
    for g in dgs:
        g.close()


def test_peel_batch_shared_launches(ctx):
    """ga_peel_run_batch (every round's launches cover all unfinished states) equals each state's own fused peel and
    the oracle: best density, best round, rounds and every vertex's removal round, at two eps, twice in a row (the
    second on fewer states), under mixed tunings and a capped grid; a wrong round-0 count is reported."""
    graphs = mix(30, 33, nmax=3000)
    graphs.append(np.array(clique(40), np.uint32))
    oracles = [oracle.OGraph.from_pairs(p) for p in graphs]
    dgs = [ga.DeviceGraph.from_pairs(ctx, p) for p in graphs]
    for i, g in enumerate(dgs):
        if i % 3 == 1:
            t = ga.TcTuning.from_dict(ga.default_tuning().as_dict())
            t.lanes[:] = [4, 8, 16, 32, 64] if i % 2 else [1, 32, 16, 64, 64]
            t.hash_cap = 8 if i % 4 == 1 else 4096   # small cap: the binary-search bucket
            g.set_tuning(t)
    Ts = [og.triangles() for og in oracles]
    peels = [ga.Peel(g) for g in dgs]
    for rep, (num, den, sel) in enumerate(((1, 10, list(range(len(dgs)))), (1, 20, list(range(0, len(dgs), 2))))):
        res, ms = ga.peel_batch(ctx, [peels[i] for i in sel], num, den, [Ts[i] for i in sel])
        assert ms > 0
        for i, r in zip(sel, res):
            bn, bd, br, rounds, rr = oracles[i].batch_peel(num, den)
            assert (r.best_num, r.best_den, r.best_round, r.rounds) == (bn, bd, br, rounds), (rep, i)
            assert np.array_equal(r.removed_round, rr), (rep, i)
            alone = ga.Peel(dgs[i])
            assert alone.run(num, den, T_full=Ts[i]).same(r), (rep, i)
            alone.close()
    ctx.set_grid_cap(3)
    try:
        res, _ = ga.peel_batch(ctx, peels, 1, 10)
        for i, r in enumerate(res):
            assert (r.best_num, r.best_den, r.best_round, r.rounds) == oracles[i].batch_peel(1, 10)[:4], i
    finally:
        ctx.set_grid_cap(0)
    bad = list(Ts)
    bad[3] += 1
    with pytest.raises(ga.GAError) as ei:
        ga.peel_batch(ctx, peels, 1, 10, bad)
    assert ei.value.code == ga.GA_ERR_STATE
    for p in peels:
        p.close()
    for g in dgs:
        g.close()
