"""The C++ drivers (tests/cpp: gpu_tc, gpu_densest, gpu_dsg, gpu_truss) end to end on the GPU: they parse a SNAP text
file with the reference's loader rules, run the drop-in host API over the C ABI, verify every result against the linked
CPU oracle and print the reference's CSV rows (gpu_tc.cu:14-21, gpu_densest.cu:16-25, gpu_dsg.cu:27-46; gpu_truss has
no reference driver) only when all of it matches."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from graph_analyzer_amd import gen

BIN = os.path.join(os.path.dirname(__file__), "cpp", "bin")


def run(args, timeout=240):
    r = subprocess.run([os.path.join(BIN, args[0]), *args[1:]], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=timeout)
    return r.returncode, r.stdout, r.stderr


@pytest.fixture(scope="module")
def graph_file(tmp_path_factory):
    raw = gen.powerlaw_small()
    path = tmp_path_factory.mktemp("drv") / "powerlaw_small.txt"
    with open(path, "w") as f:
        f.write("# seeded power-law graph\n% comment\n\n")
        np.savetxt(f, raw, fmt="%d", delimiter="\t")
    return str(path), oracle.OGraph.load_text(str(path))


def rows(out):
    return [line.split(",") for line in out.strip().splitlines()]


def test_driver_usage_errors():
    """Exit 2 and no rows on usage errors (checked before any device work, so this runs without a GPU)."""
    for args in (["gpu_tc"], ["gpu_densest"], ["gpu_densest", "g.txt", "1"], ["gpu_dsg"], ["gpu_dsg", "a.txt", "b.txt"],
                 ["gpu_dsg", "--T", "3"], ["gpu_truss"]):
        rc, out, _ = run(args)
        assert rc == 2 and out == "", args


@pytest.mark.gpu
def test_gpu_tc_driver(graph_file):
    path, og = graph_file
    rc, out, err = run(["gpu_tc", path, "2", "--no-burnin"])
    assert rc == 0, err[-2000:]
    rs = rows(out)
    assert len(rs) >= 7 and all(len(r) == 8 for r in rs)
    T = og.triangles()
    for r in rs:
        assert r[1] == path and (int(r[2]), int(r[3]), int(r[4])) == (og.n, og.m, T), r


@pytest.mark.gpu
def test_gpu_densest_driver(graph_file):
    path, og = graph_file
    rc, out, err = run(["gpu_densest", path, "1", "20", "--no-burnin", "--repeats", "2"])
    assert rc == 0, err[-2000:]
    bn, bd, br, rounds, _ = og.batch_peel(1, 20)
    rs = rows(out)
    assert {r[0] for r in rs} >= {"gpu_batch_peel", "gpu_batch_peel_fused"}
    for r in rs:
        assert (int(r[4]), int(r[5]), int(r[7])) == (bn, bd, rounds), r


@pytest.mark.gpu
def test_gpu_dsg_driver_with_prune(graph_file):
    path, og = graph_file
    rc, out, err = run(["gpu_dsg", path, "--no-burnin", "--repeats", "2", "--with-recount", "--verify-full"])
    assert rc == 0, err[-2000:]
    rs = {r[0]: r for r in rows(out)}
    assert set(rs) == {"cpu_core_decomp", "gpu_core_decomp", "gpu_core_decomp_fused", "gpu_core_decomp_recount", "cpu_greedy",
                       "prune_kcore", "cpu_greedy_pp_pruned", "gpu_greedy_pp_pruned", "gpu_core_exact_pruned_seeded",
                       "pipeline_gpu_core", "pipeline_cpu_core", "gpu_core_exact_full"}
    core, kmax = og.core()
    kt, kv = og.core_level_density(core, kmax)
    for name in ("cpu_core_decomp", "gpu_core_decomp", "gpu_core_decomp_fused"):
        assert (int(rs[name][4]), int(rs[name][5]), int(rs[name][11])) == (kt, kv, kmax)
    gt, gv, _, _ = og.greedy_pp(1)   # the lower bound is the denser of the kmax-core and Greedy (gpu_dsg.cu:318-330)
    assert (int(rs["cpu_greedy"][4]), int(rs["cpu_greedy"][5])) == (gt, gv)
    if gt * kv > kt * gv:
        kt, kv = gt, gv
    k = (kt + kv - 1) // kv
    sub, _, _ = og.prune_to_core(core, k)
    p = rs["prune_kcore"]
    assert (int(p[2]), int(p[3]), int(p[4]), int(p[10]), int(p[11])) == (sub.n, sub.m, sub.triangles(), k, kmax)
    gt, gv, _, _ = sub.greedy_pp(10)
    for name in ("cpu_greedy_pp_pruned", "gpu_greedy_pp_pruned"):
        r = rs[name]
        assert (int(r[2]), int(r[3]), int(r[4]), int(r[5]), int(r[10])) == (sub.n, sub.m, gt, gv, 10), r
    # CoreExact seeded with Greedy++ (the driver checked its set and, on small pruned graphs, Exact's density): exact,
    # so no lower than Greedy++, at least one flow call
    r = rs["gpu_core_exact_pruned_seeded"]
    et, ev = int(r[4]), int(r[5])
    assert (int(r[2]), int(r[3])) == (sub.n, sub.m) and et * gv >= gt * ev and int(r[10]) >= 1, r
    # the pipeline rows carry the optimum on the whole graph's n, m; --verify-full finds the same optimum unpruned
    for name in ("pipeline_gpu_core", "pipeline_cpu_core", "gpu_core_exact_full"):
        p = rs[name]
        assert (int(p[2]), int(p[3])) == (og.n, og.m) and int(p[4]) * ev == et * int(p[5]), p
    assert int(rs["pipeline_gpu_core"][10]) == int(rs["gpu_core_decomp_fused"][10])


@pytest.mark.gpu
def test_gpu_dsg_driver_no_greedy(graph_file):
    """--no-greedy: no cpu_greedy row, and the lower bound is the kmax-core density alone."""
    path, og = graph_file
    rc, out, err = run(["gpu_dsg", path, "--no-burnin", "--repeats", "1", "--no-greedy", "--no-greedypp"])
    assert rc == 0, err[-2000:]
    rs = {r[0]: r for r in rows(out)}
    assert set(rs) == {"cpu_core_decomp", "gpu_core_decomp", "gpu_core_decomp_fused", "prune_kcore", "gpu_core_exact_pruned_seeded",
                       "pipeline_gpu_core", "pipeline_cpu_core"}
    core, kmax = og.core()
    kt, kv = og.core_level_density(core, kmax)
    assert int(rs["prune_kcore"][10]) == (kt + kv - 1) // kv


@pytest.mark.gpu
def test_gpu_truss_driver(graph_file):
    path, og = graph_file
    rc, out, err = run(["gpu_truss", path, "--repeats", "1"])
    assert rc == 0, err[-2000:]
    r = rows(out)
    _, kmax = og.truss()
    assert len(r) == 1 and r[0][0] == "gpu_truss" and (int(r[0][2]), int(r[0][3]), int(r[0][4])) == (og.n, og.m, kmax)


@pytest.mark.gpu
def test_host_api_signatures(graph_file):
    """build_tasks, build_edge_order, build_undirected_csr and core_decomposition called through the drop-in headers'
    reference signatures (graph.h:138-171, densest.h:162-171, dsg.h:116-166), each equal to the CPU oracle."""
    path, og = graph_file
    rc, out, err = run(["test_host_api", path])
    assert rc == 0 and "ALL OK" in out, (out[-2000:], err[-2000:])
