"""Benchmark: exact triangle counting on one MI355X (oriented edges/s), with the HBM roofline of the counting kernels
and the reference's CPU counter timed on this machine's host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload rmat21|powerlaw_communities|powerlaw_small]

A step is one full exact count T = sum over oriented edges (u,v) of |N+(u) ∩ N+(v)| of the resident graph (CSR already
in HBM). The graph is generated from a fixed seed, built on the GPU (de-dup, orientation, sort; timed separately) and
checked byte for byte against the CPU oracle's CSR, and every step's count must equal the CPU count; otherwise the
script exits 1 without a result line.

Roofline: algorithmic bytes per count B = 4(n+1) + 4m + 4 * sum over oriented edges of min(d+(u), d+(v)) (read the CSR
once and, per edge, every element of the shorter out-list once); achieved = B / mean device time of one count (HIP
events on the library's stream); peak = 8 TB/s HBM3E. traffic = measured HBM bytes per count from the committed
rocprofv3 counter summary when one exists for the workload (else null).
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


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_workload(name):
    from graph_analyzer_amd import gen
    if name == "rmat21":
        return gen.rmat21(), {"workload": "rmat_scale21_ef8", "generator": "R-MAT (0.57,0.19,0.19,0.05), 16,777,216 samples, seed 2102"}
    if name == "powerlaw_communities":
        return gen.powerlaw_communities(), {"workload": "chung_lu_4M_communities",
                                            "generator": "Chung-Lu n=4.0M exponent 2.3 + 20k planted communities (10-200, p=0.3), seed 2103"}
    if name == "powerlaw_small":
        return gen.powerlaw_small(), {"workload": "chung_lu_8298", "generator": "Chung-Lu n=8,298 m=100,762 exponent 2.1, seed 2101"}
    raise SystemExit(f"unknown workload {name}")


def algorithmic_bytes(rowptr, col, src):
    d = np.diff(rowptr.astype(np.int64))
    mins = np.minimum(d[src], d[col]).astype(np.int64)
    n = rowptr.size - 1
    return 4 * (n + 1) + 4 * col.size + 4 * int(mins.sum()), int(mins.sum())


def load_traffic(workload):
    path = os.path.join(ROOT, "profiles", f"tc_hash_{workload}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("hbm_bytes_per_count")


def cpu_baseline(pairs, T_gpu, budget_s):
    """The reference's cpu_triangles_serial (compiled from its sources into oracle/_ref) when present, else the oracle
    restatement; run on the whole graph, serially (1 thread). Also checks the GPU count."""
    import oracle
    t0 = time.perf_counter()
    if oracle.ref_available():
        uniq = oracle.canonical_unique(pairs)
        n = int(uniq.max()) + 1 if uniq.size else 1
        g = oracle.RGraph.build(n, uniq)
        kind = "reference"
    else:
        g = oracle.OGraph.from_pairs(pairs)
        kind = "port"
    log(f"cpu graph built in {time.perf_counter() - t0:.1f} s ({kind})")
    t0 = time.perf_counter()
    T_omp = g.triangles_omp(0)
    t_omp = time.perf_counter() - t0
    import oracle as _o
    threads = _o.olib().orc_omp_max_threads()
    if T_omp != T_gpu:
        raise SystemExit(f"MISMATCH: GPU count {T_gpu} != CPU count {T_omp}")
    t0 = time.perf_counter()
    T_ser = g.triangles()
    t_ser = time.perf_counter() - t0
    assert T_ser == T_omp
    return {"kind": kind, "t_serial_s": t_ser, "t_omp_s": t_omp, "omp_threads": threads, "m": g.m}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="rmat21")
    ap.add_argument("--variant", default="hash")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="also run one plain device copy of MIB MiB (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (this build runs on a single MI355X)", file=sys.stderr)
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
    if variant == ga.TC_HASH and not args.no_tune:
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

    cpu = None
    if not args.no_cpu:
        cb = cpu_baseline(pairs, T0, 30)
        cpu = {"value": cb["m"] / cb["t_serial_s"], "unit": "oriented edges/s", "cores": 1, "kind": cb["kind"],
               "sample": f"whole graph ({cb['m']} oriented edges), cpu_triangles_serial, 1 thread, single run; "
                         f"OpenMP {cb['omp_threads']} threads: {cb['m'] / cb['t_omp_s']:.4g} edges/s",
               "omp_value": cb["m"] / cb["t_omp_s"], "omp_cores": cb["omp_threads"]}
    else:
        ref_T = og.triangles_omp(0)
        if ref_T != T0:
            raise SystemExit(f"MISMATCH: GPU count {T0} != CPU count {ref_T}")
    if args.copy_probe:
        import torch
        x = torch.ones(args.copy_probe << 18, dtype=torch.int32, device="cuda")
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        log(f"copy probe: {args.copy_probe} MiB read + {args.copy_probe} MiB written")
    traffic = load_traffic(cfg["workload"])
    out = {
        "metric": "oriented edges/s (exact triangle count, one resident graph)",
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
        "config": {**cfg, "n": g.n, "m": g.m, "max_out_degree": g.max_out_degree, "triangles": T0, "variant": ga.variant_name(variant),
                   "tuning": tune_info, "build_ms": g.build_ms, "plan_ms": prep_ms},
        "triangles_per_s": T0 / (wall / args.steps),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "algorithmic_bytes": B_alg, "sum_min_outdeg": summin, "kernel_ms_mean": mean_dev},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
