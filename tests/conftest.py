import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device 0) and the built libga_amd.so")
    config.addinivalue_line("markers", "slow: large inputs (tens of seconds)")


@pytest.fixture(scope="session")
def ctx():
    import graph_analyzer_amd as ga
    c = ga.Context(0)
    print(f"device: {c.device_info()}", flush=True)
    yield c
    c.close()


def clique(k, base=0):
    return [(base + a, base + b) for a in range(k) for b in range(a + 1, k)]


def zoo():
    """Hand-made graphs: the literal zoo of the reference's tiny suite (emu_dsg.cpp:405-436) plus closed forms."""
    z = []
    z.append(("no edges", 5, []))
    z.append(("path6", 6, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5)]))
    z.append(("star7", 7, [(0, i) for i in range(1, 7)]))
    z.append(("C5", 5, [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0)]))
    z.append(("K33", 6, [(a, c) for a in range(3) for c in range(3, 6)]))
    z.append(("single triangle + isolated", 7, [(0, 1), (1, 2), (0, 2)]))
    z.append(("triangle + pendant path", 6, [(0, 1), (1, 2), (0, 2), (2, 3), (3, 4), (4, 5)]))
    for k in (3, 4, 5, 6, 7, 9, 12):
        z.append((f"K{k}", k, clique(k)))
    z.append(("K4 + K4", 8, clique(4) + clique(4, 4)))
    z.append(("K5 + K4", 9, clique(5) + clique(4, 5)))
    z.append(("K5 + K3 + K3", 11, clique(5) + clique(3, 5) + clique(3, 8)))
    z.append(("two K5 sharing a vertex", 9, clique(5) + clique(5, 4)))
    z.append(("two K4 sharing an edge", 6, sorted(set(clique(4) + clique(4, 2)))))
    z.append(("K6 + pendant triangle", 8, clique(6) + [(5, 6), (6, 7), (5, 7)]))
    z.append(("wheel W8", 8, [e for i in range(1, 8) for e in ((0, i), (i, i % 7 + 1))]))
    z.append(("fan", 8, [(0, i) for i in range(1, 8)] + [(i, i + 1) for i in range(1, 7)]))
    z.append(("K5 minus edge", 5, [e for e in clique(5) if e != (0, 1)]))
    z.append(("bowtie", 5, [(0, 1), (1, 2), (0, 2), (2, 3), (3, 4), (2, 4)]))
    z.append(("K4 -- K5 bridge", 9, clique(4) + clique(5, 4) + [(3, 4)]))
    z.append(("greedy-gap 10", 10, [(0, 6), (1, 2), (1, 4), (1, 7), (1, 8), (2, 8), (3, 4), (3, 6), (4, 6), (6, 7), (7, 8), (7, 9)]))
    z.append(("greedy-gap 8", 8, [(0, 1), (0, 2), (0, 4), (1, 3), (1, 4), (1, 5), (1, 6), (2, 4), (3, 4), (3, 7), (5, 6)]))
    z.append(("greedy-gap 9", 9, [(0, 2), (0, 4), (1, 6), (1, 8), (2, 3), (2, 4), (2, 5), (2, 7), (2, 8), (3, 4), (3, 7), (6, 8)]))
    out = []
    for name, n, e in z:
        e = sorted({(min(a, b), max(a, b)) for a, b in e if a != b})
        out.append((name, n, np.array(e, dtype=np.uint32).reshape(-1, 2)))
    return out


def random_graph(rng, n, p, dup=0.0, loops=0):
    """ER pairs with optional duplicate reversed copies and self loops (the reference's gen_test_graph.py pattern)."""
    iu = np.triu_indices(n, 1)
    mask = rng.random(iu[0].size) < p
    a, b = iu[0][mask].astype(np.uint32), iu[1][mask].astype(np.uint32)
    pairs = np.stack([a, b], 1)
    if dup > 0 and pairs.size:
        d = pairs[rng.random(len(pairs)) < dup][:, ::-1]
        pairs = np.concatenate([pairs, d])
    if loops:
        v = rng.integers(0, n, loops).astype(np.uint32)
        pairs = np.concatenate([pairs, np.stack([v, v], 1)])
    if len(pairs):
        pairs = pairs[rng.permutation(len(pairs))]
    return pairs.reshape(-1, 2)
