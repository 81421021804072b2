"""Writes tests/golden/exact_vectors.json from the reference's OWN CoreExact and Exact (dsg.h:252-627, through
oracle/_ref/libref.so). Run in the build container, where the reference sources exist:

    make -C oracle && python tests/golden/make_exact_golden.py

Inputs: every graph of reference_vectors.json (raw pairs and the reference's own core numbers), plus seeded graphs with
planted dense blocks in several components and small dense random graphs. Outputs per graph: core_exact with Pruning1/2
(the default), without them (the test hook that starts from one triangle and makes the binary search shrink its network
over many core levels), and seeded with the reference's Greedy++ result (10 passes), as gpu_dsg.cu:348-361 runs it;
exact_densest (Alg 1 on the whole graph) for graphs of at most 1,000 vertices. Each result: tri, verts, the vertex set,
flow_calls, components_searched, shrinks.
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
from conftest import random_graph  # noqa: E402


def planted(rng, blocks, n, p_noise):
    """Dense blocks (size, edge probability) on disjoint id ranges, joined by a sparse random background."""
    pairs, base = [], 0
    for size, p in blocks:
        a, b = np.triu_indices(size, 1)
        keep = rng.random(a.size) < p
        pairs.append(np.stack([a[keep] + base, b[keep] + base], 1))
        base += size
    a, b = np.triu_indices(n, 1)
    keep = rng.random(a.size) < p_noise
    pairs.append(np.stack([a[keep], b[keep]], 1))
    return np.concatenate(pairs).astype(np.uint32)


def record(name, raw, tmpdir):
    path = os.path.join(tmpdir, "g.txt")
    with open(path, "w") as f:
        for a, b in raw:
            f.write(f"{a}\t{b}\n")
    g = oracle.RGraph.load_text(path)
    core, kmax = g.core()
    rec = {"name": name, "raw_pairs": np.asarray(raw, dtype=np.uint32).tolist(), "n": g.n, "m": g.m, "core": core.tolist(),
           "kmax": kmax}
    rec["core_exact"] = g.core_exact(core)
    rec["core_exact_no_pruning12"] = g.core_exact(core, pruning12=False)
    gt, gv, gs, _ = g.greedy_pp(10)
    rec["greedy_pp10"] = {"tri": gt, "verts": gv, "set": gs.tolist()}
    rec["core_exact_seeded"] = g.core_exact(core, seed=(gt, gv, gs)) if gv else g.core_exact(core)
    if g.n <= 1000:
        rec["exact_densest"] = g.exact_densest()
    for k in ("core_exact", "core_exact_no_pruning12", "core_exact_seeded"):
        assert rec[k]["ok"], (name, k)
    return rec


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/libref.so is not built (needs the reference sources)")
    with open(os.path.join(HERE, "reference_vectors.json")) as f:
        base = json.load(f)["graphs"]
    out = []
    with tempfile.TemporaryDirectory() as td:
        for g in base:
            raw = np.array(g["raw_pairs"], np.uint32).reshape(-1, 2)
            out.append(record(g["name"], raw, td))
        rng = np.random.default_rng(20261204)
        extra = [
            ("planted_two_blocks", planted(rng, [(14, 0.8), (11, 0.9)], 200, 0.02)),
            ("planted_three_blocks_sparse", planted(rng, [(20, 0.5), (16, 0.6), (12, 0.75)], 600, 0.004)),
            ("planted_nested", planted(rng, [(30, 0.35), (9, 1.0)], 400, 0.01)),
            ("dense_random_40", random_graph(rng, 40, 0.45, dup=0.0, loops=0)),
        ]
        for i in range(8):
            extra.append((f"small_dense_{i}", random_graph(rng, 9 + i, 0.55, dup=0.0, loops=0)))
        for name, raw in extra:
            out.append(record(name, raw, td))
    with open(os.path.join(HERE, "exact_vectors.json"), "w") as f:
        json.dump({"source": "reference core_exact / exact_densest (dsg.h:252-627) compiled by oracle/Makefile", "graphs": out}, f,
                  separators=(",", ":"))
    print(f"wrote {len(out)} graphs")


if __name__ == "__main__":
    main()
