"""Grid-shape independence: with every count, participation, support, peel and core launch capped at 1, 3 or 7
workgroups (ga_ctx_set_grid_cap), each grid-stride and item loop runs many passes, and every result must still equal
the CPU oracle bit for bit. This is the role of the reference's emulator harness, which runs every count kernel on 1-,
3- and 7-block grids (emu_main.cpp:40-57)."""
import json
import os

import numpy as np
import pytest

import graph_analyzer_amd as ga
import oracle
from conftest import clique, random_graph
from test_gpu_parity import ALL_VARIANTS, check_graph

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")
CAPS = (1, 3, 7)


def graphs_for(cap):
    from graph_analyzer_amd import gen
    rng = np.random.default_rng(700 + cap)
    with open(GOLDEN) as f:
        recs = json.load(f)["graphs"]
    recs = sorted(recs, key=lambda r: -r["T"])[:3]   # the fixtures with the most triangles
    out = [(r["name"], np.array(r["raw_pairs"], dtype=np.uint32).reshape(-1, 2)) for r in recs]
    out.append(("random", random_graph(rng, 400, 0.05, dup=0.2, loops=3)))
    out.append(("chung_lu", gen.chung_lu(4000, 30000, 2.1, 1.0, seed=70 + cap)))
    hub = [(0, v) for v in range(1, 1500)] + [(v, v + 1) for v in range(1, 1499, 2)] + clique(25, 1500)
    out.append(("hub", np.array(hub, np.uint32)))
    return out


@pytest.mark.parametrize("cap", CAPS)
def test_tiny_grids_single_graph(ctx, cap):
    """Count variants, participation on random alive masks, supports, peel (fused and host-driven) and core (fused,
    host-driven incremental, recount) on grids of `cap` workgroups."""
    ctx.set_grid_cap(cap)
    try:
        for name, raw in graphs_for(cap):
            print(f"grid cap {cap}: {name}", flush=True)
            check_graph(ctx, raw, eps=[(1, 10), (1, 1)], tag=f"cap {cap} {name}")
    finally:
        ctx.set_grid_cap(0)


@pytest.mark.parametrize("cap", CAPS)
def test_tiny_grids_count_shapes(ctx, cap):
    """The hash count's kernel forms on small grids: one lane per item, every lane-group width, the flat kernel for
    light rows, the binary-search bucket (small hash cap), the renumbered copy and the XCD-sliced item order."""
    from graph_analyzer_amd import gen
    raw = gen.chung_lu(20000, 150000, 2.05, 0.0, seed=80 + cap)
    og = oracle.OGraph.from_pairs(raw)
    T = og.triangles()
    _, tri_ref = og.participation()
    g = ga.DeviceGraph.from_pairs(ctx, raw)
    ctx.set_grid_cap(cap)
    try:
        for lanes, flat, hcap, order, relabel in (([1, 8, 32, 64, 64], (0, 64), 4096, 1, 0), ([4, 16, 16, 32, 64], (2048, 256), 4096, 3, 1),
                                                  ([1, 4, 8, 16, 32], (512, 64), 24, 2, 0), ([8, 32, 64, 64, 64], (0, 64), 3, 5, 2)):
            t = ga.TcTuning()
            t.lanes[:] = lanes
            t.block, t.work_chunk, t.load_shift, t.batches = 256, 4096, 2, 8
            t.flat_chunk, t.flat_block = flat
            t.hash_cap, t.item_order, t.relabel, t.probe = hcap, order, relabel, order % 2
            g.set_tuning(t)
            assert g.count(ga.TC_HASH)[0] == T, (cap, t.as_dict())
            assert np.array_equal(g.vertex_participation()[1], tri_ref), (cap, t.as_dict())
        for v in ALL_VARIANTS:
            assert g.count(v)[0] == T, (cap, ga.variant_name(v))
    finally:
        ctx.set_grid_cap(0)
        g.close()


@pytest.mark.parametrize("cap", CAPS)
def test_tiny_grids_batch(ctx, cap):
    """A batch on small grids: per-graph counts, the batched peel (all graphs in shared launches) and the batched core
    (per-state workgroup ranges) equal the oracle graph by graph."""
    from test_gpu_batch import mix
    graphs = mix(12, 90 + cap, nmax=1500)
    oracles = [oracle.OGraph.from_pairs(p) for p in graphs]
    b = ga.Batch(ctx, graphs)
    ctx.set_grid_cap(cap)
    try:
        T_batch, _ = b.count_all()
        for i, og in enumerate(oracles):
            assert T_batch[i] == og.triangles(), (cap, i)
        stats, rr, _ = b.peel_all(1, 10, T_full=T_batch)
        for i, og in enumerate(oracles):
            bn, bd, br, rounds, orr_ = og.batch_peel(1, 10)
            assert tuple(int(x) for x in stats[i]) == (bn, bd, br, rounds), (cap, i)
            assert np.array_equal(rr[b.voff[i]:b.voff[i + 1]], orr_), (cap, i)
    finally:
        ctx.set_grid_cap(0)
        b.close()
    dgs = [ga.DeviceGraph.from_pairs(ctx, p) for p in graphs]
    states = [ga.Core(g, True) for g in dgs]
    ctx.set_grid_cap(cap)
    try:
        res, _ = ga.core_batch(ctx, states, [og.triangles() for og in oracles])
        for (core, rounds, kmax), og in zip(res, oracles):
            oc_, okmax = og.core()
            assert np.array_equal(core, oc_) and kmax == okmax, cap
    finally:
        ctx.set_grid_cap(0)
        for s in states:
            s.close()
        for g in dgs:
            g.close()
