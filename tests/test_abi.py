"""The C-ABI library loads without a GPU and exports every symbol include/ga_amd.h declares; the host-API C++ test
and the generators run on the CPU."""
import ctypes
import os
import re
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "ga_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ga_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import graph_analyzer_amd as ga
    L = ctypes.CDLL(ga.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 40
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # the Python binding covers the same set
    ga.lib()


def test_no_gpu_fails_loudly():
    """Without a visible GPU the product raises instead of computing on the CPU."""
    import graph_analyzer_amd as ga
    try:
        import torch  # noqa: F401
    except Exception:
        pass
    r = subprocess.run(["python", "-c", "import graph_analyzer_amd as ga; ga.Context(0)"], cwd=ROOT,
                       env={**os.environ, "HIP_VISIBLE_DEVICES": "-1", "ROCR_VISIBLE_DEVICES": "-1"},
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "ga_ctx_create" in (r.stderr + r.stdout)
    assert ga.variant_name(ga.TC_HASH) == "hash_lds_bucketed"


def test_host_loops_cpp():
    exe = os.path.join(ROOT, "tests", "cpp", "bin", "test_host_loops")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout[-2000:]


def test_generators_deterministic():
    from graph_analyzer_amd import gen
    a = gen.rmat(12, 1 << 15, seed=5)
    b = gen.rmat(12, 1 << 15, seed=5)
    assert np.array_equal(a, b) and a.shape == (1 << 15, 2) and a.max() < (1 << 12)
    c = gen.chung_lu(5000, 40000, 2.2, 3.0, seed=9)
    assert np.array_equal(c, gen.chung_lu(5000, 40000, 2.2, 3.0, seed=9)) and c.max() < 5000
    e = gen.chung_lu_exact(1000, 5000, 2.1, 2.0, seed=3)
    canon = np.sort(e, axis=1)
    assert len(np.unique(canon[:, 0].astype(np.uint64) << 32 | canon[:, 1])) == 5000 and (e[:, 0] != e[:, 1]).all()
    cm = gen.communities(10000, 50, 10, 40, 0.3, seed=2)
    assert len(cm) > 0 and cm.max() < 10000


def test_powerlaw_small_shape():
    """The 8,298-vertex workload has exactly m = 100,762 unique edges and max degree near 1,000."""
    import oracle
    from graph_analyzer_amd import gen
    p = gen.powerlaw_small()
    g = oracle.OGraph.from_pairs(p)
    assert g.m == 100762 and g.n <= 8298
    r, c, s = g.arrays()
    deg = np.bincount(np.concatenate([c, s]), minlength=g.n)
    assert 500 < deg.max() < 2000


def test_serving_workload_and_summary():
    """The serving workload is seeded and shaped as specified; the summary's percentiles are computed per kind."""
    from graph_analyzer_amd import serve
    src, reqs = serve.mixed_workload(rate=2000.0, duration=1.0, seed=3, pool=8, rmat_graphs=2, rmat_scale=10)
    src2, reqs2 = serve.mixed_workload(rate=2000.0, duration=1.0, seed=3, pool=8, rmat_graphs=2, rmat_scale=10)
    assert [(r.kind, r.key, r.arrival) for r in reqs] == [(r.kind, r.key, r.arrival) for r in reqs2]
    assert 1800 < len(reqs) < 2200                                   # Poisson(2000) over 1 s
    share = np.mean([r.key[0] == "rmat" for r in reqs])
    assert 0.02 < share < 0.09
    assert {r.kind for r in reqs} == set(serve.KINDS) and all(r.den == 20 for r in reqs)
    assert np.all(np.diff([r.arrival for r in reqs]) >= 0)
    res = [serve.Result(r.rid, r.kind, r.key, (0,), r.arrival, r.arrival, r.arrival + 0.001 * (1 + r.rid % 10), 0, 1)
           for r in reqs]
    s = serve.summarize(res)
    assert s["requests"] == len(reqs) and 1.0 <= s["p50_ms"] <= s["p99_ms"] <= 10.0 + 1e-9
    assert set(s["by_kind"]) == set(serve.KINDS)
    assert np.array_equal(src(("mix", 3)), src2(("mix", 3)))


def test_default_tuning_and_argument_checks_on_host():
    """Host-only entry points: the default tuning record, and argument checks that reject before any device call."""
    import graph_analyzer_amd as ga
    t = ga.default_tuning()
    d = t.as_dict()
    assert d["lanes"][0] == 1 and all(x in (4, 8, 16, 32, 64) for x in d["lanes"][1:]), d
    assert d["item_order"] == 1 and d["block"] % 64 == 0 and 1 <= d["load_shift"] <= 3
    assert ga.TcTuning.from_dict(d).as_dict() == d
    L = ga.lib()
    for ctx, nbytes, seg in ((None, 1 << 20, 64), (None, 1 << 20, 48), (None, 16, 64)):
        assert L.ga_ctx_gather_probe(ctx, nbytes, seg, None, None) == ga.GA_ERR_ARG
    assert L.ga_ctx_copy_probe(None, 1 << 20, None) == ga.GA_ERR_ARG
    assert L.ga_tc_default_tuning(None) == ga.GA_ERR_ARG


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus 2 exits with status 2, a message and no result line (checked before any device
    work, so this runs without a GPU)."""
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 2 and r.stdout == "" and "--gpus 2 is not supported" in r.stderr
