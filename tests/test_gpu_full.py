"""Full-size parity: the R-MAT scale-21 graph and the 34.7M-edge power-law graph, each result bit-exact against the
CPU oracle (OpenMP where the oracle has it). Each test runs well under a minute on a GPU host."""
import numpy as np
import pytest

import graph_analyzer_amd as ga
import oracle
from graph_analyzer_amd import gen

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

_cache = {}


def _graph(ctx, name):
    if name not in _cache:
        raw = gen.rmat21() if name == "rmat21" else gen.powerlaw_communities()
        og = oracle.OGraph.from_pairs(raw)
        _cache[name] = (raw, og, og.triangles_omp(0))
    raw, og, T = _cache[name]
    g = ga.DeviceGraph.from_pairs(ctx, raw)
    r, c, s = g.download()
    orr, oc, os_ = og.arrays()
    assert np.array_equal(r, orr) and np.array_equal(c, oc) and np.array_equal(s, os_), f"{name}: CSR differs"
    assert g.count(ga.TC_HASH)[0] == T and g.count(ga.TC_MERGE)[0] == T
    return g, og, T


ALL_VARIANTS = [ga.TC_MERGE, ga.TC_GROUP4, ga.TC_GROUP8, ga.TC_GROUP16, ga.TC_GROUP64, ga.TC_HASH, ga.TC_SORTED, ga.TC_NOSRC,
                ga.TC_VERTEX]


def test_full_size_rmat21_every_variant(ctx):
    """The R-MAT configuration asks for every intersection variant bit-exact: each one counts the full R-MAT scale-21
    graph to the oracle's T."""
    g, og, T = _graph(ctx, "rmat21")
    for v in ALL_VARIANTS:
        Tv, ms = g.count(v)
        print(f"rmat21 {ga.variant_name(v)}: {ms:.2f} ms", flush=True)
        assert Tv == T, ga.variant_name(v)
    g.close()


@pytest.mark.parametrize("name", ["rmat21", "powerlaw_communities"])
def test_full_size_undirected_csr(ctx, name):
    """build_undirected_csr (densest.h:162-171) on the device, byte for byte against the oracle's."""
    g, og, T = _graph(ctx, name)
    ur, uc = g.undirected()
    our, ouc = og.undirected()
    assert np.array_equal(ur, our) and np.array_equal(uc, ouc)
    g.close()


@pytest.mark.parametrize("name", ["rmat21", "powerlaw_communities"])
def test_full_size_peel(ctx, name):
    """Batch peeling at eps 1/10 and 1/100 (the R-MAT configuration's two epsilons)."""
    g, og, T = _graph(ctx, name)
    for num, den in ((1, 10), (1, 100)):
        bn, bd, br, rounds, rr = og.batch_peel(num, den, omp=True)
        pr = g.batch_peel(num, den, T_full=T)
        assert (pr.best_num, pr.best_den, pr.best_round, pr.rounds) == (bn, bd, br, rounds), (name, num, den)
        assert np.array_equal(pr.removed_round, rr), (name, num, den)
    g.close()


@pytest.mark.parametrize("name", ["rmat21", "powerlaw_communities"])
def test_full_size_support_and_participation(ctx, name):
    g, og, T = _graph(ctx, name)
    Ts, sup = og.edge_support()
    Tg, supg = g.edge_support(ga.TC_HASH)
    assert Tg == Ts == T and np.array_equal(supg, sup)
    _, tri = og.participation(omp=True)
    Tv, trig = g.vertex_participation(None, ga.TC_HASH)
    assert Tv == T and np.array_equal(trig, tri)
    g.close()


@pytest.mark.parametrize("name", ["rmat21", "powerlaw_communities"])
def test_full_size_ingest(ctx, name, tmp_path):
    """load_graph with the text parse on the GPU, and the binary CSR file, on the full-size graphs."""
    g, og, T = _graph(ctx, name)
    path = str(tmp_path / "g.txt")
    gen.write_text(_cache[name][0], path, name)
    t = ga.DeviceGraph.load_text(ctx, path)
    for a, b in zip(t.download(), og.arrays()):
        assert np.array_equal(a, b)
    print(f"{name}: text parse + build {t.build_ms:.2f} ms", flush=True)
    bpath = str(tmp_path / "g.gacsr")
    t.save(bpath)
    h = ga.DeviceGraph.load(ctx, bpath)
    for a, b in zip(h.download(), og.arrays()):
        assert np.array_equal(a, b)
    assert h.count(ga.TC_HASH)[0] == T
    print(f"{name}: binary load + checks {h.build_ms:.2f} ms", flush=True)
    for x in (t, h, g):
        x.close()


def test_full_size_truss(ctx):
    """k-truss of the 34.7M-edge graph through properties (the oracle's exact peeling takes minutes at this size; exact
    parity runs on graphs of up to about a million edges in tests/test_gpu_truss.py): kmax equals the oracle's result
    for this seeded graph, 2 <= truss <= support + 2 for every edge, and at several levels k every edge of the edges
    with truss >= k closes at least k - 2 triangles among them (the defining property, recounted on the device)."""
    g, og, T = _graph(ctx, "powerlaw_communities")
    t, rounds, kmax = g.truss_decomposition()
    assert kmax == 83   # oracle.OGraph.truss() on this graph (one host core, about 260 s)
    Ts, sup = g.edge_support(ga.TC_HASH)
    assert Ts == T and t.min() >= 2 and np.all(t.astype(np.int64) <= sup.astype(np.int64) + 2)
    r, c, s = g.download()
    for k in sorted({3, 4, int(np.median(t)) + 1, kmax // 2, kmax - 1, kmax}):
        keep = t >= k
        h = ga.DeviceGraph.from_pairs(ctx, np.stack([s[keep], c[keep]], 1), n=g.n)
        _, hs = h.edge_support(ga.TC_HASH)
        assert h.m == int(keep.sum()) and int(hs.min()) + 2 >= k, k
        h.close()
    print(f"truss: {rounds} rounds, kmax {kmax}", flush=True)
    g.close()


def test_full_size_core(ctx):
    """(k,triangle)-core numbers of the 34.7M-edge graph against Alg 3 (dsg.h:116-166)."""
    g, og, T = _graph(ctx, "powerlaw_communities")
    core, kmax = og.core()
    c = ga.Core(g, True)
    cg, rounds, kg = c.run(T_full=T)
    assert kg == kmax and np.array_equal(cg, core)
    # prune_to_core at ceil(kmax-core density) on the device core numbers (gpu_dsg.cu:318-328)
    kt, kv = og.core_level_density(core, kmax)
    k = (kt + kv - 1) // kv
    sub, old, cout = c.prune(k)
    osub, oold, ocout = og.prune_to_core(core, k)
    assert np.array_equal(old, oold) and np.array_equal(cout, ocout)
    r, col, src = sub.download()
    orr, oc, os_ = osub.arrays()
    assert np.array_equal(r, orr) and np.array_equal(col, oc) and np.array_equal(src, os_)
    print(f"prune at k={k}: n {g.n} -> {sub.n}, m {g.m} -> {sub.m} in {sub.build_ms:.2f} ms", flush=True)
    sub.close()
    c.close()
    g.close()
