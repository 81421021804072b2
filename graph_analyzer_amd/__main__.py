"""Runs one stage of the path on a named synthetic workload and prints its device time and result.

    python -m graph_analyzer_amd count|peel|core|support|truss|ingest|prune|greedy|exact
                                 [--workload rmat21|powerlaw_communities|powerlaw_small] [--repeat R] [--num 1 --den 10]
                                 [--passes 10]
    python -m graph_analyzer_amd batch [--graphs 4096] [--repeat R] [--tune]
    python -m graph_analyzer_amd greedy_many [--graphs 256] [--passes 10]   (seeded dense graphs of 100-1,000 vertices)
    python -m graph_analyzer_amd probe     (copy bandwidth and random-segment gather rates of the device)
    python -m graph_analyzer_amd serve [--rate 2000] [--seconds 1] [--pool 256]
                                 (continuous-batching server on the mixed request workload: throughput and latency)
    python -m graph_analyzer_amd core_batch [--rate 2000] [--seconds 1] [--pool 256]
                                 (every core request of that mixed workload in one ga_core_run_batch: device time)

Meant for profiling a single stage (rocprofv3 -- python3 -m graph_analyzer_amd core --workload rmat21) and for quick
checks on a GPU host; results are not compared with a CPU oracle here (the tests and bench.py do that).
"""
from __future__ import annotations

import argparse
import sys
import time

import numpy as np

import graph_analyzer_amd as ga
from graph_analyzer_amd import gen

WORKLOADS = {"rmat21": gen.rmat21, "powerlaw_communities": gen.powerlaw_communities, "powerlaw_small": gen.powerlaw_small}


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m graph_analyzer_amd")
    ap.add_argument("stage", choices=["count", "peel", "core", "support", "truss", "ingest", "prune", "greedy", "exact", "greedy_many", "batch",
                                      "probe", "serve", "core_batch"])
    ap.add_argument("--workload", default="rmat21", choices=sorted(WORKLOADS))
    ap.add_argument("--repeat", type=int, default=3)
    ap.add_argument("--num", type=int, default=1)
    ap.add_argument("--den", type=int, default=10)
    ap.add_argument("--graphs", type=int, default=4096)
    ap.add_argument("--passes", type=int, default=10)
    ap.add_argument("--rate", type=float, default=2000.0)
    ap.add_argument("--seconds", type=float, default=1.0)
    ap.add_argument("--pool", type=int, default=256)
    ap.add_argument("--tune", action="store_true", help="batch: autotune the batched count first (as bench.py does)")
    ap.add_argument("--slots", type=int, default=4, help="serve: worker streams")
    ap.add_argument("--core-slots", type=int, default=None, help="serve: workers that may run core work at once (default half of them)")
    args = ap.parse_args(argv)
    if args.stage == "serve":   # one Poisson run of the mixed workload through the continuous-batching server
        from graph_analyzer_amd import serve
        source, reqs = serve.mixed_workload(args.rate, args.seconds, pool=args.pool)
        s = serve.Server(source, slots=args.slots, count_batch=64, core_batch=1024, peel_batch=1024, states_per_graph=16,
                         core_slots=args.core_slots)
        t0 = time.perf_counter()
        s.preload(sorted({q.key for q in reqs}))
        print(f"serve: {len(reqs)} requests at {args.rate:g} req/s, graphs resident in {time.perf_counter() - t0:.1f} s", flush=True)
        for _ in range(args.repeat):
            summ = serve.summarize(s.run(reqs))
            print(f"serve: {summ['throughput_rps']:.0f} req/s, p50 {summ['p50_ms']:.1f} ms, p99 {summ['p99_ms']:.1f} ms, "
                  f"makespan {summ['makespan_s']:.2f} s, by kind {summ['by_kind']}", flush=True)
        s.close()
        return 0
    ctx = ga.Context(0)
    if args.stage == "core_batch":   # the serving mix's core requests as one shared-launch batch (the server's core path)
        from graph_analyzer_amd import serve
        source, reqs = serve.mixed_workload(args.rate, args.seconds, pool=args.pool)
        cores = [q for q in reqs if q.kind == "core"]
        dgs = {k: ga.DeviceGraph.from_pairs(ctx, source(k)) for k in sorted({q.key for q in cores})}
        Ts = {k: g.count(ga.TC_HASH)[0] for k, g in dgs.items()}
        states = [ga.Core(dgs[q.key], True) for q in cores]
        print(f"core_batch: {len(cores)} core requests on {len(dgs)} graphs "
              f"({sum(1 for q in cores if q.key[0] == 'rmat')} on R-MAT graphs)", flush=True)
        for _ in range(args.repeat):
            res, ms = ga.core_batch(ctx, states, [Ts[q.key] for q in cores])
            print(f"core_batch: {ms:.1f} ms device, rounds max {max(r[1] for r in res)}", flush=True)
        for st in states:
            st.close()
        for g in dgs.values():
            g.close()
        return 0
    if args.stage == "probe":
        for _ in range(2):
            ms = ctx.copy_probe(1 << 30)
            print(f"copy 1 GiB: {ms:.3f} ms = {2 * (1 << 30) / ms / 1e9:.2f} TB/s read + write", flush=True)
        for mib in (139, 2048):
            for seg in (32, 64, 128, 256):
                best = None
                for _ in range(3):
                    n, ms = ctx.gather_probe(mib << 20, seg)
                    best = ms if best is None else min(best, ms)
                lines = n * max(seg // 128, 1)
                print(f"gather {mib} MiB, {seg}-byte segments: {best:.3f} ms, {n / best / 1e6:.1f} G segments/s, "
                      f"{lines / best / 1e6:.1f} G 128-byte lines/s, {n * seg / best / 1e9:.2f} TB/s useful", flush=True)
        return 0
    if args.stage == "greedy_many":   # Greedy++ on many small dense graphs: one launch vs one call per graph
        rng = np.random.default_rng(2024)
        graphs = []
        for _ in range(args.graphs):
            n = int(np.exp(rng.uniform(np.log(100), np.log(1000))))
            p = float(rng.uniform(0.05, 0.4))
            a, b = np.triu_indices(n, 1)
            keep = rng.random(a.size) < p
            graphs.append(ga.DeviceGraph.from_pairs(ctx, np.stack([a[keep], b[keep]], 1).astype(np.uint32)))
        print(f"greedy_many: {len(graphs)} graphs, n {sum(g.n for g in graphs)}, m {sum(g.m for g in graphs)}", flush=True)
        for _ in range(args.repeat):
            st, _, ms = ctx.greedy_pp_many(graphs, args.passes)
            one = sum(g.greedy_pp(args.passes)[4] for g in graphs)
            print(f"greedy++ {args.passes} passes: one launch {ms:.2f} ms, one call per graph {one:.2f} ms (device), "
                  f"{one / ms:.1f}x", flush=True)
        for g in graphs:
            g.close()
        return 0
    if args.stage == "batch":
        bp = gen.BatchPairs(gen.batch_mix_specs(args.graphs))
        b = ga.Batch(ctx, flat=bp.pairs, offsets=bp.offsets)
        del bp
        print(f"batch {args.graphs} graphs n={b.total_n} m={b.total_m} build {b.build_ms:.1f} ms", flush=True)
        if args.tune:
            t, tms = b.tune()
            print(f"batch tuned: {t.as_dict()} ({tms:.2f} ms)", flush=True)
        for _ in range(args.repeat):
            T, ms = b.count_all()
            print(f"batched count: {ms:.3f} ms, T total {int(T.sum())}", flush=True)
        st, _, ms = b.peel_all(args.num, args.den, T_full=T, removed_round=False)
        print(f"batched peel {args.num}/{args.den}: {ms:.3f} ms, rounds max {int(st[:, 3].max())}", flush=True)
        b.close()
        return 0
    if args.stage == "ingest":   # load_graph with the text parse on the GPU, then the binary CSR file
        import os
        import tempfile
        with tempfile.TemporaryDirectory() as td:
            path, bpath = os.path.join(td, "g.txt"), os.path.join(td, "g.gacsr")
            gen.write_text(WORKLOADS[args.workload](), path, args.workload)
            print(f"{args.workload}: {os.path.getsize(path)} bytes of text", flush=True)
            for _ in range(args.repeat):
                t0 = time.perf_counter()
                g = ga.DeviceGraph.load_text(ctx, path)
                print(f"load_text: n={g.n} m={g.m} {g.build_ms:.3f} ms device, {(time.perf_counter() - t0) * 1e3:.1f} ms wall", flush=True)
                g.save(bpath)
                g.close()
                t0 = time.perf_counter()
                h = ga.DeviceGraph.load(ctx, bpath)
                print(f"binary load: {h.build_ms:.3f} ms device, {(time.perf_counter() - t0) * 1e3:.1f} ms wall", flush=True)
                h.close()
        return 0
    t0 = time.perf_counter()
    g = ga.DeviceGraph.from_pairs(ctx, WORKLOADS[args.workload]())
    print(f"{args.workload}: n={g.n} m={g.m} max_out={g.max_out_degree} build {g.build_ms:.2f} ms "
          f"(generated and built in {time.perf_counter() - t0:.1f} s)", flush=True)
    T, ms = g.count(ga.TC_HASH)
    print(f"count: T={T} {ms:.3f} ms (first, includes nothing but the count)", flush=True)
    sub = None
    sub_core = None
    if args.stage in ("greedy", "exact"):   # Greedy++ / CoreExact on the graph prune_to_core leaves at ceil(kmax-core density)
        c = ga.Core(g, True)
        core, _, kmax = c.run(T_full=T)
        Tk, _ = g.vertex_participation((core >= kmax).astype(np.uint8))
        verts = int((core >= kmax).sum())
        sub, _, sub_core = c.prune((Tk + verts - 1) // verts if verts else 0)
        c.close()
        print(f"{args.stage} on the pruned graph: n={sub.n} m={sub.m}", flush=True)
    for _ in range(args.repeat):
        if args.stage == "greedy":
            t, v, s, _, ms = sub.greedy_pp(args.passes)
            print(f"greedy++ {args.passes} passes: {t}/{v} ({len(s)} vertices) {ms:.3f} ms device", flush=True)
        elif args.stage == "exact":   # CoreExact seeded with Greedy++ (gpu_dsg.cu:348-361), wall time
            t, v, s, _, _ = sub.greedy_pp(args.passes)
            t0 = time.perf_counter()
            ex = sub.core_exact(sub_core, seed=(t, v, s))
            print(f"core_exact seeded with {t}/{v}: {ex['tri']}/{ex['verts']}, {ex['flow_calls']} flow calls, "
                  f"{(time.perf_counter() - t0) * 1e3:.1f} ms wall", flush=True)
        elif args.stage == "count":
            T, ms = g.count(ga.TC_HASH)
            print(f"count: T={T} {ms:.3f} ms", flush=True)
        elif args.stage == "support":
            t0 = time.perf_counter()
            Ts, sup = g.edge_support(ga.TC_HASH)
            print(f"support: sum={int(sup.astype(np.uint64).sum())} (= 3T: {int(sup.astype(np.uint64).sum()) == 3 * T}) "
                  f"{(time.perf_counter() - t0) * 1e3:.1f} ms wall", flush=True)
        elif args.stage == "peel":
            p = ga.Peel(g)
            r = p.run(args.num, args.den, T_full=T)
            print(f"peel {args.num}/{args.den}: best {r.best_num}/{r.best_den} round {r.best_round} of {r.rounds}, "
                  f"{p.last_ms:.3f} ms device", flush=True)
            p.close()
        elif args.stage == "truss":
            tr = ga.Truss(g)
            t, rounds, kmax = tr.run()
            print(f"truss: rounds {rounds} kmax {kmax} {tr.last_ms:.3f} ms device", flush=True)
            tr.close()
        elif args.stage == "prune":   # prune_to_core at ceil(kmax-core density) on the device core numbers
            c = ga.Core(g, True)
            core, rounds, kmax = c.run(T_full=T)
            Tk, _ = g.vertex_participation((core >= kmax).astype(np.uint8))
            verts = int((core >= kmax).sum())
            k = (Tk + verts - 1) // verts if verts else 0
            sub, _, _ = c.prune(k)
            print(f"prune k={k}: n {g.n} -> {sub.n}, m {g.m} -> {sub.m}, {sub.build_ms:.3f} ms device", flush=True)
            sub.close()
            c.close()
        else:
            c = ga.Core(g, True)
            core, rounds, kmax = c.run(T_full=T)
            print(f"core: rounds {rounds} kmax {kmax} {c.last_ms:.3f} ms device", flush=True)
            c.close()
    if sub is not None:
        sub.close()
    g.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
