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


# SNAP text corner cases for load_graph's fgets(512) + strtoul loop (graph.h:71-89); shared with the GPU ingest tests.
# Ids that cast to 0xFFFFFFFF (a '-1', or any overflow, which strtoul saturates to ULONG_MAX) are not among them: the
# reference computes n = max_id + 1 in uint32, which wraps to 0 and is undefined there; the GPU loader rejects them.
# Cast ids are paired with a larger plain id, because the reference takes max_id from the larger endpoint before the
# cast, which can leave a vertex id >= n otherwise.
TEXT_CASES = {
    "empty": b"",
    "comments only": b"# a\n% b\n\n",
    "no final newline": b"1 2\n2 3\n3 1",
    "crlf and tabs": b"1\t2\r\n2 \t 3\r\n\r\n3\t1\r\n",
    "leading space, signs": b"  4 5\n+6 +7\n\t8\t9\n+4 -0\n - 3\n-18446744073709551615 3\n",
    "one number, junk": b"7\n8 x\nx 8\n9 10 11\n12 13junk\n#14 15\n %16 17\n",
    "loops and duplicates": b"5 5\n1 2\n2 1\n1 2\n",
    "ids beyond 2^32 (cast)": b"4294967297 3\n4294967300 2\n18446744069414584321 5\n0 9\n",
    "nul bytes": b"1 2\n3\x00 4\n5 \x006\n\x007 8\n9 10\n",
    "long line split by fgets": b"1 2" + b" " * 600 + b"3 4\n5 6\n",
    "number cut at byte 511": b" " * 509 + b"12 34\n1 2\n",
    "pair cut at byte 511": b" " * 505 + b"21 43 65 87\n",
    "line of exactly 511 and 512 bytes": b"1 2" + b" " * 507 + b"\n" + b"3 4" + b" " * 508 + b"\n9 8\n",
    "very long lines": (b"7 9 " + b"0" * 1500 + b"11 12 \n") * 3 + b"#" + b" " * 700 + b"\n",
}


# Texts on which this library's loader departs from the reference on purpose, with the result it gives: a pair that
# equals another only after the cast to uint32 ("1 3" and "4294967297 3", which casts to (3, 1)). The reference sorts
# and de-duplicates the cast pairs without putting them back in (min, max) order, so it keeps (1, 3) and (3, 1) and
# hands the builder a repeated edge (m = 3 here, a multigraph); the GPU loader and the oracle canonicalise after the
# cast and keep one (n, m) = (10, 2).
DIVERGENT_TEXT_CASES = {
    "pair equal to another after the cast": (b"1 3\n4294967297 3\n0 9\n", 10, 2),
}
