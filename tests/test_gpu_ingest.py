"""GPU ingest: load_graph's text parse on the device (ga_graph_from_text / ga_graph_load_text) and the binary CSR file
(ga_graph_save / ga_graph_load). The text path must give byte-identical CSRs to the CPU oracle's loader (itself pinned
against the reference's own load_graph on the same corner cases) and to the host parse followed by the GPU builder; the
binary path must round-trip exactly and reject every malformed file."""
import ctypes

import numpy as np
import pytest

import graph_analyzer_amd as ga
import oracle
from conftest import DIVERGENT_TEXT_CASES, TEXT_CASES, random_graph
from graph_analyzer_amd import gen

pytestmark = pytest.mark.gpu


def same_csr(dg, og, tag):
    r, c, s = dg.download()
    orr, oc, os_ = og.arrays()
    assert (dg.n, dg.m, dg.max_out_degree) == (og.n, og.m, og.max_out_degree), tag
    assert np.array_equal(r, orr) and np.array_equal(c, oc) and np.array_equal(s, os_), f"CSR differs: {tag}"


def test_text_corner_cases(ctx, tmp_path):
    for name, text in TEXT_CASES.items():
        path = tmp_path / "g.txt"
        path.write_bytes(text)
        og = oracle.OGraph.load_text(str(path))
        same_csr(ga.DeviceGraph.from_text(ctx, text), og, name)
        same_csr(ga.DeviceGraph.load_text(ctx, str(path)), og, name + " (file)")


def test_text_divergent_cases(ctx, tmp_path):
    """Inputs where the loader departs from the reference on purpose (conftest.DIVERGENT_TEXT_CASES): the GPU gives
    the oracle's documented result."""
    for name, (text, n, m) in DIVERGENT_TEXT_CASES.items():
        path = tmp_path / "g.txt"
        path.write_bytes(text)
        og = oracle.OGraph.load_text(str(path))
        assert (og.n, og.m) == (n, m), name
        same_csr(ga.DeviceGraph.from_text(ctx, text), og, name)
        same_csr(ga.DeviceGraph.load_text(ctx, str(path)), og, name + " (file)")


def test_text_ids_that_do_not_fit(ctx):
    """A '-1' or an overflowing number casts to 0xFFFFFFFF, so n = max_id + 1 does not fit uint32: rejected."""
    for text in (b"-1 5\n", b"1 2\n99999999999999999999999 6\n", b"4294967295 0\n"):
        with pytest.raises(ga.GAError) as ei:
            ga.DeviceGraph.from_text(ctx, text)
        assert ei.value.code == ga.GA_ERR_ARG, text


def _format(rng, pairs):
    """Random SNAP formatting: separators, CRLF, comment and blank lines, leading space, '+' signs, a final line with no
    newline, and a few lines longer than fgets' 511 bytes."""
    seps = [b" ", b"\t", b"  ", b" \t "]
    out = [b"# generated\n", b"% other comment\n"]
    for i, (a, b) in enumerate(pairs.tolist()):
        r = rng.random()
        lead = b"  " if r < 0.05 else (b"+" if r < 0.08 else b"")
        end = b"\r\n" if r > 0.9 else b"\n"
        line = lead + str(a).encode() + seps[int(rng.integers(4))] + str(b).encode()
        if r > 0.995:
            line += b" " * int(rng.integers(505, 1600)) + b"7 8"   # continues as a second fgets chunk
        out.append(line + end)
        if rng.random() < 0.01:
            out.append(b"\n" if rng.random() < 0.5 else b"# note\n")
    text = b"".join(out)
    return text[:-1] if text.endswith(b"\n") else text


@pytest.mark.parametrize("seed", range(3))
def test_text_random_formatting(ctx, tmp_path, seed):
    rng = np.random.default_rng(40 + seed)
    raw = random_graph(rng, int(rng.integers(200, 4000)), float(rng.uniform(0.001, 0.02)), dup=0.3, loops=5)
    if seed == 2:
        raw = np.concatenate([raw, gen.powerlaw_small()])
    text = _format(rng, raw)
    path = tmp_path / "g.txt"
    path.write_bytes(text)
    og = oracle.OGraph.load_text(str(path))
    g = ga.DeviceGraph.from_text(ctx, text)
    same_csr(g, og, seed)
    assert g.count(ga.TC_HASH)[0] == og.triangles()
    g.close()


def test_binary_round_trip(ctx, tmp_path):
    for raw in (gen.powerlaw_small(), np.zeros((0, 2), np.uint32), np.array([[0, 1]], np.uint32),
                random_graph(np.random.default_rng(3), 900, 0.02, dup=0.2, loops=2)):
        g = ga.DeviceGraph.from_pairs(ctx, raw)
        path = str(tmp_path / "g.gacsr")
        g.save(path)
        h = ga.DeviceGraph.load(ctx, path)
        assert (h.n, h.m, h.max_out_degree) == (g.n, g.m, g.max_out_degree)
        for a, b in zip(h.download(), g.download()):
            assert np.array_equal(a, b)
        if g.m:
            assert h.count(ga.TC_HASH)[0] == g.count(ga.TC_HASH)[0]
        g.close()
        h.close()
    # a graph with no vertices (an empty induced subgraph) round-trips too
    g = ga.DeviceGraph.from_pairs(ctx, gen.powerlaw_small())
    e, _ = g.induced(np.zeros(g.n, np.uint8))
    path = str(tmp_path / "empty.gacsr")
    e.save(path)
    h = ga.DeviceGraph.load(ctx, path)
    assert (h.n, h.m, h.max_out_degree) == (0, 0, 0) and h.download()[0].tolist() == [0]
    for x in (e, h, g):
        x.close()


def test_binary_rejects_malformed(ctx, tmp_path):
    g = ga.DeviceGraph.from_pairs(ctx, gen.powerlaw_small())
    good = str(tmp_path / "good.gacsr")
    g.save(good)
    data = bytearray(open(good, "rb").read())
    n, m = g.n, g.m
    hdr = 64
    rowptr = np.frombuffer(bytes(data[hdr: hdr + 4 * (n + 1)]), np.uint32)
    col_off = hdr + 4 * (n + 1)
    u = int(np.flatnonzero(np.diff(rowptr.astype(np.int64)) >= 2)[0])     # a row with two entries
    e = int(rowptr[u])
    bad = {}
    b = bytearray(data); b[0:8] = b"GACSR999"; bad["magic"] = b
    bad["truncated"] = data[:-4]
    b = bytearray(data); b[col_off + 4 * e: col_off + 4 * e + 4] = np.uint32(n).tobytes(); bad["col >= n"] = b
    c0 = data[col_off + 4 * e: col_off + 4 * e + 4]
    c1 = data[col_off + 4 * e + 4: col_off + 4 * e + 8]
    b = bytearray(data); b[col_off + 4 * e: col_off + 4 * e + 8] = c1 + c0; bad["row not ascending"] = b
    b = bytearray(data); b[hdr + 4 * n: hdr + 4 * n + 4] = np.uint32(m - 1).tobytes(); bad["rowptr[n] != m"] = b
    b = bytearray(data); b[24:32] = np.uint64(g.max_out_degree + 1).tobytes(); bad["max_out_degree"] = b
    # reversing one edge u -> v into v -> u keeps every row sorted only if v's row is rebuilt, so instead orient every
    # edge of the file the wrong way: the CSR of the reversed edges, built from the pairs, then written raw
    r, c, s = g.download()
    rev = np.stack([c, s], 1)
    order = np.lexsort((rev[:, 1], rev[:, 0]))
    rev = rev[order]
    rrow = np.zeros(n + 1, np.uint32)
    np.add.at(rrow, rev[:, 0].astype(np.int64) + 1, 1)
    rrow = np.cumsum(rrow).astype(np.uint32)
    maxo = int(np.diff(rrow.astype(np.int64)).max())
    head = bytearray(data[:hdr]); head[24:32] = np.uint64(maxo).tobytes()
    bad["reversed orientation"] = head + rrow.tobytes() + rev[:, 1].astype(np.uint32).tobytes()
    for name, blob in bad.items():
        path = tmp_path / "bad.gacsr"
        path.write_bytes(bytes(blob))
        with pytest.raises(ga.GAError) as ei:
            ga.DeviceGraph.load(ctx, str(path))
        assert ei.value.code == ga.GA_ERR_ARG, name
    with pytest.raises(ga.GAError):
        ga.DeviceGraph.load(ctx, str(tmp_path / "missing.gacsr"))
    out = ga._vp()
    assert ga.lib().ga_graph_from_text(ctx.handle, None, 5, ctypes.byref(out), None) == ga.GA_ERR_ARG and not out
    g.close()
