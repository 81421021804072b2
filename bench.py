"""Benchmark: exact triangle counting on one MI355X (oriented edges/s and the HBM roofline of the count), plus the
4,096-graph batch (graphs/s, one launch set for all graphs vs one count per graph), with the CPU oracle's restatement of
the reference's CPU code (count, batch peel, Alg 3 core) timed on this machine's host cores (model, nproc and OpenMP
threads printed).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload powerlaw_communities|rmat21|powerlaw_small]
                    [--batch-graphs 4096] [--serve-seconds 1.0] [--serve-rates 250,500,1000,2000] [--no-cpu] [--no-tune]

Count step: one full exact count T = sum over oriented edges (u,v) of |N+(u) ∩ N+(v)| of the resident graph (CSR
already in HBM). The graph is generated from a fixed seed, built on the GPU (de-dup, orientation, sort; timed
separately) and checked byte for byte against the CPU oracle's CSR, and every step's count must equal the CPU count;
otherwise the script exits 1 without a result line.

Roofline: algorithmic bytes per count B = 4(n+1) + 4m + 4 * sum over oriented edges of min(d+(u), d+(v)) (read the CSR
once and, per edge, every element of the shorter out-list once); achieved = B / mean device time of one count (HIP
events on the library's stream around the count's kernels); peak = 8 TB/s HBM3E. traffic = HBM bytes per count
measured with rocprofv3 counters (FETCH_SIZE + WRITE_SIZE of the count kernels, scaled by the same counters on a plain
copy of known size, --copy-probe) when profiles/tc_hash_<workload>_traffic.json holds them, else null.

Batch: the 4,096-graph mix (n log-uniform in [1k, 200k], mean degree uniform in [4, 64], half Erdos-Renyi, half
Chung-Lu) resident as one disjoint union; graphs/s = graphs / wall time of one per-graph count of every graph, batched
(one launch set) and one count per graph (same kernels, each graph its own launches and readback; the baseline is the
faster of the library's default tuning and the batch's tuning). Both must give identical per-graph counts; a sample of
graphs is also checked against the CPU oracle. The batch roofline: B_alg summed over the graphs over the device time of
the batched launch set. Peel: batch peeling at eps = 1/10, all graphs' rounds in shared launches vs one fused peel per
graph.

CPU rows (same workload): the count serial and with OpenMP (the cpu_baseline object), batch peeling at eps = 1/10 with
OpenMP and serially, and serial Alg 3, each beside the GPU time and each checked against the GPU result; the serial
peel and core run in a background thread while the GPU legs run.

Extras (same workload): load_graph with the text parse on the GPU and the binary CSR file, and prune_to_core at
ceil(kmax-core density) on the device core numbers, each checked against the oracle and timed beside the oracle's CPU
restatement on one host core.

Serving: a load sweep of Poisson arrivals (250, 500, 1,000 and 2,000 req/s, 1 s each by default) of count / peel
(eps 1/20) / core requests, 95% on a pool of 256 graphs of the batch size law and 5% on two R-MAT scale-20 graphs, all
resident; at each rate, p50 / p99 latency (arrival to result) and throughput of the continuous-batching server vs the
plain one-request-at-a-time form on the same arrivals, whose results must be identical.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


_T0 = time.perf_counter()


def log(msg):
    print(f"[bench {time.perf_counter() - _T0:6.1f} s] {msg}", file=sys.stderr, flush=True)


def make_workload(name):
    from graph_analyzer_amd import gen
    if name == "rmat21":
        return gen.rmat21(), {"workload": "rmat_scale21_ef8", "generator": "R-MAT (0.57,0.19,0.19,0.05), 16,777,216 samples, seed 2102"}
    if name == "powerlaw_communities":
        return gen.powerlaw_communities(), {"workload": "chung_lu_4M_communities",
                                            "generator": "Chung-Lu n=4.0M exponent 2.3, 15.2M samples + 20k planted communities "
                                                         "(10-200 vertices, p=0.3), seed 2103"}
    if name == "powerlaw_small":
        return gen.powerlaw_small(), {"workload": "chung_lu_8298", "generator": "Chung-Lu n=8,298 m=100,762 exponent 2.1, seed 2101"}
    raise SystemExit(f"unknown workload {name}")


def algorithmic_bytes(rowptr, col, src):
    d = np.diff(rowptr.astype(np.int64))
    mins = np.minimum(d[src], d[col]).astype(np.int64)
    n = rowptr.size - 1
    return 4 * (n + 1) + 4 * col.size + 4 * int(mins.sum()), int(mins.sum())


def bucket_rooflines(g, B_alg, reps=20):
    """Roofline of each launch of one count: the launch's share of B_alg (4 bytes per probe element of its entries, plus
    the CSR structure term 4(n+1) + 4m split over the plan's entries) over its mean device time (HIP events between the
    launches). The shares add up to B_alg."""
    ent, el, ms = g.bucket_profile(reps)
    lanes = list(g.get_tuning().lanes)
    names = ["d+(h)<=8", "d+(h)<=32", "d+(h)<=128", "d+(h)<=512", "d+(h)<=4096", "binary search (d+(h) above the LDS cap)", "flat kernel (light rows)"]
    flat = ms[6] > 0
    if flat:   # the flat kernel takes the light rows (buckets 0-2)
        ent[6], el[6] = sum(ent[:3]), sum(el[:3])
        ent[:3], el[:3] = [0, 0, 0], [0, 0, 0]
    nw = sum(ent)
    struct = B_alg - 4 * sum(el)   # 4(n+1) + 4m
    out = []
    for i in range(7):
        if not ent[i]:
            continue
        b = 4 * el[i] + struct * ent[i] / nw
        ach = b / (ms[i] * 1e-3) / 1e9 if ms[i] > 0 else 0.0
        out.append({"launch": names[i], "lanes": (lanes[i] if i < 5 else (64 if i == 5 else None)), "entries": int(ent[i]),
                    "probe_elements": int(el[i]), "algorithmic_bytes": b, "ms": ms[i], "achieved": ach, "frac": ach / HBM_PEAK_GBS})
    return out


def load_traffic(workload):
    path = os.path.join(ROOT, "profiles", f"tc_hash_{workload}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("hbm_bytes_per_count")


def host_info():
    """CPU model, logical CPUs of this process and OpenMP's thread count (the CPU baseline's cores)."""
    import oracle
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"cpu_model": model, "nproc": usable, "os_cpu_count": os.cpu_count(), "omp_max_threads": int(oracle.olib().orc_omp_max_threads())}


def cpu_baseline(og, T_gpu):
    """The CPU oracle's restatement of the reference's counters (cpu_triangles_serial / cpu_triangles_omp,
    graph.h:110-136: one thread, then OpenMP schedule(dynamic,64) on every host core), on the whole graph; each run must
    reproduce the GPU count. Best of 3 for the OpenMP run (cpu_tc.cpp:15-19); the serial one runs once (8-10 s)."""
    import oracle
    threads = int(oracle.olib().orc_omp_max_threads())
    t_omp = None
    for _ in range(3):
        t0 = time.perf_counter()
        T_omp = og.triangles_omp(0)
        t = time.perf_counter() - t0
        t_omp = t if t_omp is None else min(t_omp, t)
        if T_omp != T_gpu:
            raise SystemExit(f"MISMATCH: GPU count {T_gpu} != CPU count {T_omp}")
    t0 = time.perf_counter()
    T_ser = og.triangles()
    t_ser = time.perf_counter() - t0
    if T_ser != T_omp:
        raise SystemExit(f"MISMATCH: serial CPU count {T_ser} != OpenMP count {T_omp}")
    return {"kind": "port", "t_serial_s": t_ser, "t_omp_s": t_omp, "omp_threads": threads, "m": og.m}


def cpu_peel_core_rows(og, gpu_peel, start_serial):
    """CPU rows beside the GPU peel (eps = 1/10) and core on the same graph (densest.h:138-160, dsg.h:125-166, as
    cpu_densest.cpp:27-38 runs them, single runs): batch peeling with OpenMP on every host core now, then, in a
    background thread (ctypes releases the GIL), serial batch peeling and serial Alg 3 while the GPU legs run; each
    result must equal the GPU's. Returns (rows, join) where join() waits for the serial rows and adds them."""
    import threading
    rows = {}
    t0 = time.perf_counter()
    r = og.batch_peel(1, 10, omp=True)
    rows["peel_omp_s"] = time.perf_counter() - t0
    if r[:4] != (gpu_peel.best_num, gpu_peel.best_den, gpu_peel.best_round, gpu_peel.rounds) or not np.array_equal(r[4], gpu_peel.removed_round):
        raise SystemExit("MISMATCH: GPU batch peel differs from the CPU's")
    err = []

    def serial():
        try:
            t1 = time.perf_counter()
            rs = og.batch_peel(1, 10, omp=False)
            rows["peel_serial_s"] = time.perf_counter() - t1
            if rs[:4] != r[:4]:
                err.append("serial and OpenMP CPU peels differ")
            t1 = time.perf_counter()
            core, kmax = og.core()
            rows["core_serial_s"] = time.perf_counter() - t1
            rows["core_kmax"] = int(kmax)
            rows["_core"] = core
        except Exception as e:   # reported by join()
            err.append(repr(e))

    th = threading.Thread(target=serial, daemon=True)
    if start_serial:
        th.start()

    def join(gpu_core):
        if not start_serial:
            return rows
        th.join()
        if err:
            raise SystemExit("MISMATCH: " + "; ".join(err))
        core = rows.pop("_core")
        if gpu_core is not None and not np.array_equal(core, gpu_core):
            raise SystemExit("MISMATCH: GPU core numbers differ from CPU Alg 3")
        return rows

    return rows, join


def run_extras(ctx, g, pairs, cfg, csr, T0):
    """The rows next to the count on the same workload, each checked against the CPU oracle and timed beside the
    oracle's restatement of the reference's CPU code, 1 thread:
    GPU text ingest (load_graph, graph.h:69-96) and the binary CSR file; GPU prune_to_core (dsg.h:636-645) at
    ceil(kmax-core density) on the device core numbers; Greedy++ (dsg.h:189-248) with 10 passes on the pruned graph;
    the k-truss decomposition (GPU time, checked against the per-edge supports)."""
    import tempfile

    import graph_analyzer_amd as ga
    import oracle
    from graph_analyzer_amd import gen
    out = {"cpu_kind": "port", "cpu_cores": 1}
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "g.txt")
        gen.write_text(pairs, path, cfg["workload"])
        size = os.path.getsize(path)
        walls, devs = [], []
        for _ in range(3):
            t0 = time.perf_counter()
            t = ga.DeviceGraph.load_text(ctx, path)
            walls.append((time.perf_counter() - t0) * 1e3)
            devs.append(t.build_ms)
            if len(walls) == 1 and not all(np.array_equal(a, b) for a, b in zip(t.download(), csr)):
                raise SystemExit("MISMATCH: GPU text ingest differs from the oracle's CSR")
            if len(walls) < 3:
                t.close()
        t0 = time.perf_counter()
        rg = oracle.OGraph.load_text(path)
        t_cpu = time.perf_counter() - t0
        if not all(np.array_equal(a, b) for a, b in zip(rg.arrays(), csr)):
            raise SystemExit("MISMATCH: CPU load_graph differs from the oracle's CSR")
        del rg
        bpath = os.path.join(td, "g.gacsr")
        t.save(bpath)
        t.close()
        bwalls, bdevs = [], []
        for _ in range(3):
            t0 = time.perf_counter()
            h = ga.DeviceGraph.load(ctx, bpath)
            bwalls.append((time.perf_counter() - t0) * 1e3)
            bdevs.append(h.build_ms)
            h.close()
        out["ingest"] = {"text_bytes": size, "lines": int(len(pairs)) + 1,
                         "gpu_parse_build_ms": min(devs), "gpu_wall_ms": min(walls),
                         "gpu_text_gb_per_s": size / (min(devs) * 1e-3) / 1e9,
                         "cpu_load_graph_s": t_cpu, "speedup_wall": t_cpu * 1e3 / min(walls),
                         "binary_bytes": os.path.getsize(bpath), "binary_load_wall_ms": min(bwalls), "binary_load_device_ms": min(bdevs)}
    log(f"ingest: {size / 1e6:.0f} MB text, GPU parse+build {min(devs):.1f} ms device / {min(walls):.1f} ms wall "
        f"(file read + upload); CPU load_graph {t_cpu:.2f} s; binary load {min(bwalls):.1f} ms wall")
    _, sup_full = g.edge_support(ga.TC_HASH)
    c = ga.Core(g, True)
    core, rounds, kmax = c.run(T_full=T0)
    T_k, _ = g.vertex_participation((core >= kmax).astype(np.uint8))
    verts = int((core >= kmax).sum())
    k = (T_k + verts - 1) // verts if verts else 0
    pms = []
    for _ in range(3):
        sub, old, cout = c.prune(k)
        pms.append(sub.build_ms)
        if len(pms) < 3:
            sub.close()
    cg = oracle.OGraph.from_csr(csr[0], csr[1])
    t0 = time.perf_counter()
    osub, oold, ocout = cg.prune_to_core(core, k)
    t_cpu = time.perf_counter() - t0
    if not (np.array_equal(old, oold) and np.array_equal(cout, ocout) and
            all(np.array_equal(a, b) for a, b in zip(sub.download(), osub.arrays()))):
        raise SystemExit("MISMATCH: GPU prune_to_core differs from the CPU's")
    out["prune"] = {"k": int(k), "kmax": int(kmax), "core_rounds": int(rounds), "core_ms": c.last_ms, "n_in": g.n, "m_in": g.m,
                    "n_out": sub.n, "m_out": sub.m, "gpu_ms": min(pms), "cpu_prune_to_core_s": t_cpu}
    log(f"prune: k={k} n {g.n} -> {sub.n}, m {g.m} -> {sub.m}: GPU {min(pms):.2f} ms, CPU prune_to_core {t_cpu:.2f} s")
    # Greedy++ with 10 passes on the pruned graph (gpu_dsg.cu:344-356): device time, best of 3, against the CPU's
    passes = 10
    if sub.n <= ctx.greedy_max_vertices():
        gms = []
        for _ in range(3):
            gt, gv, gs, gpp, ms = sub.greedy_pp(passes)
            gms.append(ms)
        t0 = time.perf_counter()
        ct, cv, cs, cpp = osub.greedy_pp(passes)
        t_g = time.perf_counter() - t0
        if (gt, gv, gpp) != (ct, cv, cpp) or not np.array_equal(gs, cs):
            raise SystemExit("MISMATCH: GPU greedy_pp differs from the CPU's")
        out["greedy_pp"] = {"passes": passes, "n": sub.n, "m": sub.m, "tri": gt, "verts": gv, "gpu_ms": min(gms), "cpu_greedy_pp_s": t_g}
        log(f"greedy++: {passes} passes on the pruned graph: {gt}/{gv}, GPU {min(gms):.2f} ms, CPU greedy_pp {t_g:.2f} s")
        # CoreExact on the pruned graph seeded with Greedy++ (gpu_dsg.cu:348-361): triangle passes on the device,
        # max-flow on the host; wall time. Checked: the set holds the claimed triangles (oracle count) and the density is
        # no lower than Greedy++'s
        t0 = time.perf_counter()
        ex = sub.core_exact(cout, seed=(gt, gv, gs))
        t_ex = time.perf_counter() - t0
        keep = np.zeros(sub.n, np.uint8)
        keep[ex["set"]] = 1
        if osub.participation(keep)[0] != ex["tri"] or len(ex["set"]) != ex["verts"] or ex["tri"] * gv < gt * ex["verts"]:
            raise SystemExit("MISMATCH: core_exact's set fails the oracle's check")
        out["core_exact"] = {"n": sub.n, "m": sub.m, "tri": ex["tri"], "verts": ex["verts"], "flow_calls": ex["flow_calls"],
                             "wall_s": t_ex, "seed": [gt, gv]}
        log(f"core_exact (seeded) on the pruned graph: {ex['tri']}/{ex['verts']}, {ex['flow_calls']} flow calls, {t_ex:.2f} s wall")
    sub.close()
    c.close()
    # k-truss: device time of the whole decomposition (support count included); best of 2. No CPU leg here: the
    # oracle's sequential peeling needs minutes on this graph.
    tr = ga.Truss(g)
    tms = []
    for _ in range(2):
        t, rounds, kmax_t = tr.run()
        tms.append(tr.last_ms)
    tr.close()
    if t.min() < 2 or np.any(t.astype(np.int64) > np.asarray(sup_full, np.int64) + 2):
        raise SystemExit("MISMATCH: a truss number outside [2, support + 2]")
    out["truss"] = {"gpu_ms": min(tms), "rounds": rounds, "kmax": kmax_t, "edges_per_s": g.m / (min(tms) * 1e-3)}
    log(f"truss: {rounds} rounds, kmax {kmax_t}, GPU {min(tms):.1f} ms")
    return out, core


def run_batch(ctx, count, steps, tune):
    """Batched vs one-count-per-graph on the batch mix; returns the result object or exits on any mismatch."""
    import graph_analyzer_amd as ga
    import oracle
    from graph_analyzer_amd import gen
    t0 = time.perf_counter()
    specs = gen.batch_mix_specs(count)
    bp = gen.BatchPairs(specs)
    log(f"batch: generated {count} graphs, {bp.pairs.shape[0]} raw pairs in {time.perf_counter() - t0:.1f} s")
    b = ga.Batch(ctx, flat=bp.pairs, offsets=bp.offsets)
    del bp
    log(f"batch: union n={b.total_n} m={b.total_m} built in {b.build_ms:.0f} ms")
    # per-graph baseline with the library's default tuning (the first pass builds the per-graph plans)
    Tp_def, _ = b.count_all(per_graph=True)
    per_default = min(b.count_all(per_graph=True)[1] for _ in range(2))
    tune_info = None
    log(f"batch: per-graph count with the default tuning {per_default:.1f} ms; tuning the batched count")
    if tune:
        t, tms = b.tune()
        tune_info = t.as_dict()
        log(f"batch: tuned {tune_info} ({tms:.2f} ms)")
    T0, _ = b.count_all()
    times_b, dev_b = [], []
    for _ in range(steps):
        T, ms = b.count_all()
        if not np.array_equal(T, T0):
            raise SystemExit("MISMATCH: batched per-graph counts changed between steps")
        times_b.append(ms)
        dev_b.append(b.work()[1])
    summin, _ = b.work()
    B_alg = 4 * (int(b.total_n) + count) + 4 * int(b.total_m) + 4 * int(summin)
    # Per-graph baseline: the fastest of the library's default tuning, the batch's tuning (inherited by every graph),
    # and each of those with the flat kernel in small chunks (one launch for all light rows of a small graph, the
    # per-graph form's best shapes on this mix)
    tuned = b.get_tuning()
    per_by = {"default": per_default}
    cands = [("batch", tuned)]
    for name, base_t in (("default", ga.default_tuning()), ("batch", tuned)):
        for fc, fb in ((512, 64), (1024, 128)):
            t2 = ga.TcTuning.from_dict(base_t.as_dict())
            t2.flat_chunk, t2.flat_block = fc, fb
            cands.append((f"{name}+flat{fc}/{fb}", t2))
    # each also with its bucket launches overlapped on four streams (the per-graph path keeps its best tuning)
    for name, base_t in [("default", ga.default_tuning())] + list(cands):
        t2 = ga.TcTuning.from_dict(base_t.as_dict())
        t2.streams = 1 - base_t.streams
        cands.append((f"{name}+streams{t2.streams}", t2))
    for name, t2 in cands:
        log(f"batch: per-graph count with the {name} tuning")
        b.set_tuning(t2)
        Tp, _ = b.count_all(per_graph=True)
        if not np.array_equal(Tp, T0):
            raise SystemExit(f"MISMATCH: per-graph launches ({name}) disagree with the batched counts")
        per_by[name] = min(b.count_all(per_graph=True)[1] for _ in range(2))
    b.set_tuning(tuned)
    if not np.array_equal(Tp_def, T0):
        raise SystemExit("MISMATCH: per-graph launches disagree with the batched counts")
    per = list(per_by.values())
    log("batch: per-graph count ms " + ", ".join(f"{k} {v:.1f}" for k, v in per_by.items()))
    st, rr, pms = b.peel_all(1, 10, T_full=T0)
    st2, rr2, pms2 = b.peel_all(1, 10, T_full=T0, per_graph=True)
    if not (np.array_equal(st, st2) and np.array_equal(rr, rr2)):
        raise SystemExit("MISMATCH: batched peel disagrees with per-graph peel")
    # CPU oracle on a sample of graphs (smallest, largest, and a spread in between)
    order = np.argsort(b.m)
    sample = sorted(set(int(order[k]) for k in np.linspace(0, count - 1, 8).astype(int)))
    for i in sample:
        r, c, _ = b.graph(i)
        og = oracle.OGraph.from_csr(r, c)
        bn, bd, br, rounds, rr_o = og.batch_peel(1, 10, omp=True)
        if og.triangles_omp(0) != int(T0[i]) or (int(st[i, 0]), int(st[i, 1]), int(st[i, 2]), int(st[i, 3])) != (bn, bd, br, rounds):
            raise SystemExit(f"MISMATCH: batch graph {i} differs from the CPU oracle")
        v0 = int(b.voff[i])
        if not np.array_equal(rr[v0: v0 + int(b.n[i])], rr_o):
            raise SystemExit(f"MISMATCH: batch graph {i} removal rounds differ from the CPU oracle")
    t_b, t_p = float(np.mean(times_b)), float(min(per))
    res = {"workload": f"batch_mix_{count}", "graphs": count, "total_n": int(b.total_n), "total_m": int(b.total_m),
           "triangles": int(T0.sum()), "tuning": tune_info,
           "count_graphs_per_s": count / (t_b * 1e-3), "count_ms": t_b,
           "per_graph_count_graphs_per_s": count / (t_p * 1e-3), "per_graph_count_ms": t_p, "count_speedup": t_p / t_b,
           "per_graph_count_ms_by_tuning": per_by, "per_graph_count_baseline": min(per_by, key=per_by.get),
           "count_edges_per_s": b.total_m / (t_b * 1e-3),
           "roofline": {"bound": "hbm", "algorithmic_bytes": B_alg, "sum_min_outdeg": int(summin),
                        "count_device_ms_mean": float(np.mean(dev_b)),
                        "achieved": B_alg / (float(np.mean(dev_b)) * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": B_alg / (float(np.mean(dev_b)) * 1e-3) / 1e9 / HBM_PEAK_GBS,
                        "note": "B_alg summed over the 4,096 graphs; device time of the batched launch set (HIP events)"},
           "peel_eps": "1/10", "peel_graphs_per_s": count / (pms * 1e-3), "peel_ms": pms,
           "per_graph_peel_graphs_per_s": count / (pms2 * 1e-3), "per_graph_peel_ms": pms2, "peel_speedup": pms2 / pms,
           "oracle_checked_graphs": sample}
    b.close()
    return res


def run_serving(rates, duration):
    """Load sweep: at each offered rate, the same Poisson arrivals served by continuous batching (4 worker streams,
    shared-launch count, peel and core batches, footprint admission) and by the plain form (one request at a time, one
    count per launch, one decomposition at a time); every result must be identical in both forms."""
    from graph_analyzer_amd import serve
    source = serve.workload_source(seed=2106, pool=256, rmat_scale=20)
    sweep = {rate: serve.poisson_requests(rate, duration, seed=2106, pool=256) for rate in rates}
    keys = sorted({q.key for reqs in sweep.values() for q in reqs})
    out = {"workload": f"poisson_{duration:g}s_per_rate", "duration_s": duration,
           "mix": "count / peel eps 1/20 / core uniform; 95% on 256 graphs of the batch size law, 5% on 2 R-MAT scale-20 graphs",
           "sweep": []}
    forms = (("continuous", dict(slots=4, count_batch=64, core_batch=1024, peel_batch=1024, states_per_graph=16)),
             ("plain", dict(slots=1, count_batch=1, core_batch=1, peel_batch=1, states_per_graph=1)))
    rows = {rate: {"offered_rps": rate, "requests": len(sweep[rate])} for rate in rates}
    values = {}
    for name, kw in forms:
        s = serve.Server(source, **kw)
        t0 = time.perf_counter()
        s.preload(keys)
        log(f"serving ({name}): {len(keys)} graphs resident in {time.perf_counter() - t0:.1f} s, {s.used / 1e9:.1f} GB")
        for rate in rates:
            res = s.run(sweep[rate])
            summ = serve.summarize(res)
            values[(name, rate)] = [r.value for r in res]
            rows[rate][name] = {k: summ[k] for k in ("throughput_rps", "p50_ms", "p99_ms", "makespan_s", "by_kind")}
            log(f"serving ({name}) at {rate:g} req/s: {summ['throughput_rps']:.0f} req/s, p50 {summ['p50_ms']:.1f} ms, "
                f"p99 {summ['p99_ms']:.1f} ms, makespan {summ['makespan_s']:.2f} s")
        s.close()
    for rate in rates:
        if values[("continuous", rate)] != values[("plain", rate)]:
            raise SystemExit(f"MISMATCH: a serving result at {rate:g} req/s differs between the continuous and the plain runs")
        out["sweep"].append(rows[rate])
    out["results_identical_alone_and_in_mix"] = True
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="powerlaw_communities")
    ap.add_argument("--variant", default="hash")
    ap.add_argument("--batch-graphs", type=int, default=4096, help="graphs in the batch comparison (0: skip it)")
    ap.add_argument("--serve-seconds", type=float, default=1.0, help="seconds of Poisson arrivals in the serving comparison (0: skip)")
    ap.add_argument("--serve-rates", default="250,500,1000,2000", help="offered requests per second of the serving load sweep")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the ingest and prune measurements")
    ap.add_argument("--tuning-from", metavar="JSON",
                    help="apply a recorded count tuning (no tuner): a result line of this script or a JSON object with a 'tuning' key")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="also run one plain device copy of MIB MiB (k_copy_probe; calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the library runs on a single MI355X; sharding is a design, not code)", file=sys.stderr)
        sys.exit(2)

    import graph_analyzer_amd as ga
    import oracle

    ctx = ga.Context(0)
    info = ctx.device_info()
    log(f"device {info}")
    t0 = time.perf_counter()
    pairs, cfg = make_workload(args.workload)
    log(f"generated {len(pairs)} raw pairs in {time.perf_counter() - t0:.1f} s")
    g = ga.DeviceGraph.from_pairs(ctx, pairs)
    log(f"GPU CSR build: n={g.n} m={g.m} max_out={g.max_out_degree} in {g.build_ms:.2f} ms (device)")
    rowptr, col, src = g.download()
    t0 = time.perf_counter()
    og = oracle.OGraph.from_pairs(pairs)
    orr, oc, os_ = og.arrays()
    if not (np.array_equal(rowptr, orr) and np.array_equal(col, oc) and np.array_equal(src, os_)):
        raise SystemExit("MISMATCH: GPU-built CSR differs from the CPU oracle's")
    log(f"CSR verified against the oracle ({time.perf_counter() - t0:.1f} s)")
    B_alg, summin = algorithmic_bytes(rowptr, col, src)

    variant = ga.TC_VARIANTS[args.variant]
    prep_ms = g.prepare() if variant == ga.TC_HASH else 0.0
    tune_info = None
    if variant == ga.TC_HASH and args.tuning_from:
        with open(args.tuning_from) as f:
            text = f.read().strip()
        try:
            rec = json.loads(text)["tuning"]
        except (ValueError, KeyError):
            rec = json.loads(text.splitlines()[-1])["config"]["tuning"]
        t = g.get_tuning()
        for k, v in rec.items():
            if k == "lanes":
                t.lanes[:] = v
            else:
                setattr(t, k, v)
        g.set_tuning(t)
        tune_info = t.as_dict()
        log(f"tuning from {args.tuning_from}: {tune_info}")
    elif variant == ga.TC_HASH and not args.no_tune:
        t, tms = g.tune()
        tune_info = t.as_dict()
        log(f"tuned: {tune_info} best {tms:.3f} ms")
    T0, _ = g.count(variant)
    for _ in range(args.warmup):
        g.count(variant)
    ctx.sync()
    dev_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        T, ms = g.count(variant)
        dev_ms.append(ms)
        if T != T0:
            raise SystemExit(f"MISMATCH: count changed between steps ({T} vs {T0})")
    ctx.sync()
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3
    mean_dev = float(np.mean(dev_ms))
    value = g.m / (wall / args.steps)
    achieved = B_alg / (mean_dev * 1e-3) / 1e9
    log(f"T={T0} ms/step={ms_per_step:.4f} device mean {mean_dev:.4f} ms (min {min(dev_ms):.4f})")
    per_launch = bucket_rooflines(g, B_alg) if variant == ga.TC_HASH else None
    for r in per_launch or []:
        log(f"  launch {r['launch']} ({r['lanes']} lanes): {r['entries']} entries, B {r['algorithmic_bytes'] / 1e9:.3f} GB, "
            f"{r['ms']:.4f} ms -> {r['achieved']:.0f} GB/s = {100 * r['frac']:.1f} %")
    if args.copy_probe:
        cms = ctx.copy_probe(args.copy_probe << 20)
        log(f"copy probe: {args.copy_probe} MiB read + {args.copy_probe} MiB written in {cms:.3f} ms")

    # GPU batch peeling at eps = 1/10 on the same graph (device time of the fused round loop), for the CPU rows
    pms = []
    for _ in range(3):
        p = ga.Peel(g)
        pr = p.run(1, 10, T_full=T0)
        pms.append(p.last_ms)
        p.close()
    gpu_peel = {"eps": "1/10", "gpu_ms": min(pms), "rounds": pr.rounds, "best": [pr.best_num, pr.best_den]}
    log(f"peel eps 1/10: {pr.rounds} rounds, best {pr.best_num}/{pr.best_den}, GPU {min(pms):.2f} ms")
    cpu = None
    join_serial = None
    if not args.no_cpu:
        hi = host_info()
        cb = cpu_baseline(og, T0)
        rows, join_serial = cpu_peel_core_rows(og, pr, start_serial=True)
        cpu = {"value": cb["m"] / cb["t_serial_s"], "unit": "oriented edges/s", "cores": 1, "kind": cb["kind"],
               "sample": f"whole graph ({cb['m']} oriented edges): the oracle's restatement of cpu_triangles_serial, 1 thread, "
                         f"single run ({cb['t_serial_s']:.2f} s); OpenMP schedule(dynamic,64) on {cb['omp_threads']} threads, "
                         f"best of 3: {cb['m'] / cb['t_omp_s']:.4g} edges/s ({cb['t_omp_s']:.2f} s)",
               "omp_value": cb["m"] / cb["t_omp_s"], "omp_cores": cb["omp_threads"], "host": hi,
               "peel_eps_1_10": {"omp_s": rows["peel_omp_s"], "omp_cores": cb["omp_threads"], "gpu_ms": gpu_peel["gpu_ms"]}}
        log(f"cpu: {hi}; count serial {cb['t_serial_s']:.2f} s, OpenMP {cb['t_omp_s']:.2f} s; peel OpenMP {rows['peel_omp_s']:.2f} s")
    else:
        ref_T = og.triangles_omp(0)
        if ref_T != T0:
            raise SystemExit(f"MISMATCH: GPU count {T0} != CPU count {ref_T}")
    traffic = load_traffic(cfg["workload"])
    n, m, max_out = g.n, g.m, g.max_out_degree
    extras = None
    gpu_core = None
    if not args.no_extras:
        extras, gpu_core = run_extras(ctx, g, pairs, cfg, (rowptr, col, src), T0)
    batch = None
    if args.batch_graphs:
        g.close()
        batch = run_batch(ctx, args.batch_graphs, max(3, min(args.steps, 10)), not args.no_tune)
        log(f"batch: {batch['count_graphs_per_s']:.0f} graphs/s batched vs {batch['per_graph_count_graphs_per_s']:.0f} per graph "
            f"({batch['count_speedup']:.2f}x); peel {batch['peel_speedup']:.2f}x")
    if join_serial is not None:
        t0 = time.perf_counter()
        rows = join_serial(gpu_core)
        del og
        cpu["peel_eps_1_10"].update({"serial_s": rows["peel_serial_s"], "cores_serial": 1})
        cpu["core"] = {"serial_s": rows["core_serial_s"], "kmax": rows["core_kmax"], "cores": 1,
                       "gpu_ms": extras["prune"]["core_ms"] if extras else None,
                       "note": "serial Alg 3 (dsg.h:125-166); no OpenMP form exists in the reference"}
        cpu["serial_rows_note"] = "serial peel and core ran in a background thread while the GPU legs ran"
        log(f"cpu: peel serial {rows['peel_serial_s']:.1f} s, core serial {rows['core_serial_s']:.1f} s "
            f"(waited {time.perf_counter() - t0:.1f} s for them)")
    else:
        del og
    serving = None
    if args.serve_seconds > 0:
        serving = run_serving([float(x) for x in args.serve_rates.split(",")], args.serve_seconds)
    out = {
        "metric": "oriented edges/s (exact triangle count, one resident graph); batched graphs/s vs per-graph launch",
        "value": value,
        "unit": "oriented edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 ids, u64 counts",
        "data": "synthetic (seeded generator)",
        "config": {**cfg, "n": n, "m": m, "max_out_degree": max_out, "triangles": T0, "variant": ga.variant_name(variant),
                   "tuning": tune_info, "build_ms": g.build_ms, "plan_ms": prep_ms},
        "triangles_per_s": T0 / (wall / args.steps),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "algorithmic_bytes": B_alg, "sum_min_outdeg": summin, "kernel_ms_mean": mean_dev,
                     "per_launch": per_launch},
        "cpu_baseline": cpu,
        "peel": gpu_peel,
        "extras": extras,
        "batch": batch,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
