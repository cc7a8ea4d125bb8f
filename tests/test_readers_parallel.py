"""The chunked, multi-threaded parse of the edge-file readers.

Both reference readers carry the two ids of a line into the next one when a
line does not write them (shared/graph.cpp:50-57, src/graph.cpp:82-90), so a
parse that cuts the file into chunks must replay each chunk's first lines
with the state the previous chunk leaves.  These tests cut quirky files into
chunks of a few bytes (KC_PARSE_CHUNK) and require the result of the one-pass
parse; the one-pass result itself is pinned against the reference readers
compiled from their sources (oracle/_ref) where those are built.  No GPU."""
import os

import numpy as np
import pytest

import oracle
import pico_amd as pa

HEADER = "%%MatrixMarket matrix coordinate integer symmetric\n100 100 100\n"


def _quirky(rng, n_lines, lo, hi, kinds):
    """Lines drawn from `kinds`: full, pair, single id, blank, text, loop,
    spaced, crlf."""
    out = []
    for _ in range(n_lines):
        k = kinds[int(rng.integers(len(kinds)))]
        a, b = (int(x) for x in rng.integers(lo, hi, 2))
        out.append({
            "full": "%d %d 1" % (a, b),
            "pair": "%d %d" % (a, b),
            "single": "%d" % a,
            "blank": "",
            "text": "x%d %d" % (a, b),
            "loop": "%d %d 1" % (a, a),
            "spaced": "  %d \t %d   7 " % (a, b),
            "crlf": "%d %d 1\r" % (a, b),
        }[k])
    return HEADER + "\n".join(out) + ("\n" if rng.integers(2) else "")


def _read(reader, path):
    try:
        g = pa.HostGraph.read_coo(path) if reader == "coo" else pa.HostGraph.read_dedup(path)
    except pa.KCoreError as e:
        return ("error", e.status)
    return ("graph", g.V, g.edges_read, g.has_zero_id, g.rowptr, g.col)


def _same(a, b):
    if a[0] != b[0]:
        return False
    if a[0] == "error":
        return a[1] == b[1]
    return a[1:4] == b[1:4] and np.array_equal(a[4], b[4]) and np.array_equal(a[5], b[5])


COO_KINDS = ["full", "pair", "single", "blank", "text", "loop", "spaced", "crlf"]
# the dedup reader decrements carried ids on every line: keep ids well above
# the number of consecutive carried lines, and no text (0 ids are refused)
DEDUP_KINDS = ["full", "full", "full", "pair", "single", "blank", "loop", "spaced", "crlf"]


@pytest.mark.parametrize("reader", ["coo", "dedup"])
@pytest.mark.parametrize("seed", range(8))
def test_chunked_parse_equals_one_pass(monkeypatch, tmp_path, reader, seed):
    rng = np.random.default_rng(100 + seed)
    kinds = COO_KINDS if reader == "coo" else DEDUP_KINDS
    lo = 0 if (reader == "coo" and seed % 2) else 40
    path = str(tmp_path / "q.txt")
    with open(path, "w") as f:
        f.write(_quirky(rng, int(rng.integers(50, 3000)), lo, 400, kinds))
    monkeypatch.setenv("KC_PARSE_CHUNK", str(1 << 40))  # one chunk: the one-pass parse
    want = _read(reader, path)
    for threads in ("1", "3", "8"):
        monkeypatch.setenv("KC_HOST_THREADS", threads)
        for chunk in (1, 5, 23, 200, 4096):
            monkeypatch.setenv("KC_PARSE_CHUNK", str(chunk))
            got = _read(reader, path)
            assert _same(got, want), (threads, chunk, got[:2], want[:2])


def test_chunk_starting_with_carried_lines(monkeypatch, tmp_path):
    # runs of blank and single-id lines straddle every chunk cut
    lines = []
    for i in range(400):
        lines += ["%d %d 1" % (i % 37 + 1, i % 53 + 2), "", "", "%d" % (i % 11 + 3), ""]
    path = str(tmp_path / "carry.txt")
    with open(path, "w") as f:
        f.write(HEADER + "\n".join(lines) + "\n")
    monkeypatch.setenv("KC_PARSE_CHUNK", str(1 << 40))
    want = _read("coo", path)
    assert want[0] == "graph"
    for chunk in (1, 2, 3, 7, 8, 9, 13):
        monkeypatch.setenv("KC_PARSE_CHUNK", str(chunk))
        assert _same(_read("coo", path), want), chunk


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (no reference tree)")
@pytest.mark.parametrize("seed", range(4))
def test_chunked_parse_matches_reference_readers(monkeypatch, tmp_path, seed):
    rng = np.random.default_rng(200 + seed)
    path = str(tmp_path / "r.txt")
    with open(path, "w") as f:
        f.write(_quirky(rng, int(rng.integers(100, 2000)), 40, 300, DEDUP_KINDS))
    monkeypatch.setenv("KC_PARSE_CHUNK", "17")
    coo = oracle.ref_read_coo(path, str(tmp_path))
    hg = pa.HostGraph.read_coo(path)
    npn, el, parts = hg.vgraph(64)
    assert hg.V == coo["num_nodes"]
    np.testing.assert_array_equal(npn, coo["node_pointer"])
    np.testing.assert_array_equal(el, coo["edge_list"])
    np.testing.assert_array_equal(parts, coo["parts"])
    csr = oracle.ref_read_csr(path, str(tmp_path))
    hd = pa.HostGraph.read_dedup(path)
    assert hd.V == csr["V"]
    np.testing.assert_array_equal(hd.rowptr.astype(np.uint32), csr["offsets"])
    np.testing.assert_array_equal(hd.col, csr["neighbors"])


@pytest.mark.parametrize("body,reader", [
    ("0 1\n2 -1\n", "coo"),            # id 2^32 - 1 beside a 0 id: V would wrap to 0
    ("1 2\n3 -1\n", "coo"),            # id 2^32 - 1, 1-based: 2^32 - 1 vertices
    ("5 4294967295 1\n", "coo"),
    ("0 4294967294 1\n", "coo"),       # 0 id + 2^32 - 2: 2^32 - 1 vertices
    ("1 99999999999 1\n", "coo"),      # overflow stores UINT_MAX
    ("5 4294967295 1\n", "dedup"),     # after the shift: 2^32 - 1 vertices
    ("1 0 1\n", "dedup"),              # 0 id wraps to UINT_MAX
])
def test_out_of_range_ids_are_refused(tmp_path, body, reader):
    path = str(tmp_path / "bad.txt")
    with open(path, "w") as f:
        f.write(HEADER + body)
    got = _read(reader, path)
    assert got == ("error", 1), got


def test_empty_and_header_only_files(tmp_path):
    for body in ("", "only one line\n", HEADER, HEADER + "\n\n"):
        path = str(tmp_path / "e.txt")
        with open(path, "w") as f:
            f.write(body)
        g = pa.HostGraph.read_coo(path)
        assert g.V == 0 and g.nnz == 0
        d = pa.HostGraph.read_dedup(path)
        assert d.V == 1 and d.nnz == 0  # V++ of an empty maximum (src/graph.cpp:102)
