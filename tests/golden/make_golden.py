"""Writes the golden fixtures of tests/golden/ from the reference's OWN CPU code (oracle/_ref/libref.so, compiled from the
reference headers by oracle/Makefile). Run in the build container, where the reference sources exist:

    make -C oracle && python tests/golden/make_golden.py

Inputs: the literal zoo of the reference's tiny suite, and seeded random graphs shaped like its emulator inputs
(Erdos-Renyi with 30% duplicate reversed edges and a self loop, gen_test_graph.py; preferential attachment,
gen_powerlaw.py), written as SNAP text so the reference's own load_graph parses them. Outputs per graph: the oriented
CSR, T, per-vertex and per-edge counts, batch-peel results for five eps values, Alg-3 core numbers with kmax and the
kmax-core density, the v9 edge order, prune_to_core (induced subgraph on core >= k) at k in {0, 1, ceil(kmax/2), kmax,
kmax + 1}, and greedy_pp with 1, 3 and 10 passes (best density, its vertex set, the best after every pass).
"""
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle  # noqa: E402
from conftest import random_graph, zoo  # noqa: E402

EPS = [(1, 10), (1, 2), (1, 1), (3, 1), (0, 1)]


def pref_attach(rng, n, k):
    edges = set()
    rep = []
    for v in range(k, n):
        chosen = set()
        while len(chosen) < k:
            chosen.add(int(rep[rng.integers(len(rep))]) if rep and rng.random() < 0.9 else int(rng.integers(v)))
        for t in chosen:
            edges.add((min(v, t), max(v, t)))
            rep += [v, t]
    return np.array(sorted(edges), dtype=np.uint32).reshape(-1, 2)


def record(name, raw_pairs, tmpdir):
    path = os.path.join(tmpdir, "g.txt")
    with open(path, "w") as f:
        f.write("# golden input\n% comment line\n\n")
        for a, b in raw_pairs:
            f.write(f"{a}\t{b}\n")
        f.write("7 x\n")   # a line with one number only: skipped by the loader
    g = oracle.RGraph.load_text(path)
    r, c, s = g.arrays()
    T = g.triangles()
    _, tri = g.participation()
    _, sup = g.edge_support()
    core, kmax = g.core()
    kt, kv = g.core_level_density(core, kmax)
    peel = {}
    for num, den in EPS:
        bn, bd, br, rounds, rr = g.batch_peel(num, den)
        peel[f"{num}/{den}"] = {"best_num": bn, "best_den": bd, "best_round": br, "rounds": rounds, "removed_round": rr.tolist()}
    prune = []
    for k in sorted({0, 1, (kmax + 1) // 2, kmax, kmax + 1}):
        pg, old, cout = g.prune_to_core(core, k)
        pr, pc, _ = pg.arrays()
        prune.append({"k": k, "n": pg.n, "m": pg.m, "max_out_degree": pg.max_out_degree, "rowptr": pr.tolist(), "col": pc.tolist(),
                      "old_id": old.tolist(), "core_out": cout.tolist()})
    greedy = {}
    for passes in (1, 3, 10):
        gt, gv, gs, gpp = g.greedy_pp(passes)
        greedy[str(passes)] = {"tri": gt, "verts": gv, "set": gs.tolist(), "per_pass": gpp}
    return {
        "name": name,
        "raw_pairs": np.asarray(raw_pairs, dtype=np.uint32).tolist(),
        "n": g.n, "m": g.m, "max_out_degree": g.max_out_degree,
        "rowptr": r.tolist(), "col": c.tolist(), "src": s.tolist(),
        "T": T, "tri": tri.tolist(), "sup": sup.tolist(),
        "core": core.tolist(), "kmax": kmax, "kmax_core": [kt, kv],
        "edge_order": g.edge_order().tolist(),
        "peel": peel,
        "prune": prune,
        "greedy_pp": greedy,
    }


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/libref.so is not built (needs the reference sources)")
    out = []
    with tempfile.TemporaryDirectory() as td:
        for name, n, e in zoo():
            raw = e.copy()
            if len(raw):
                raw = np.concatenate([raw, raw[: max(1, len(raw) // 3), ::-1]])     # duplicate reversed copies
            raw = np.concatenate([raw, np.array([[n - 1, n - 1]], np.uint32)])      # a self loop on the last id
            # the last id must appear in a real edge for n = max_id + 1 to equal the zoo's n
            rec = record(name, raw, td)
            out.append(rec)
        rng = np.random.default_rng(20261203)
        seeded = [
            ("er300", random_graph(rng, 300, 0.05, dup=0.3, loops=1)),
            ("er800", random_graph(rng, 800, 0.02, dup=0.3, loops=1)),
            ("pa400", pref_attach(rng, 400, 4)),
            ("pa900", pref_attach(rng, 900, 6)),
        ]
        for name, raw in seeded:
            out.append(record(name, raw, td))
    with open(os.path.join(HERE, "reference_vectors.json"), "w") as f:
        json.dump({"source": "reference CPU code (graph.h, densest.h, dsg.h) compiled by oracle/Makefile", "eps": EPS,
                   "graphs": out}, f, separators=(",", ":"))
    print(f"wrote {len(out)} graphs")


if __name__ == "__main__":
    main()
