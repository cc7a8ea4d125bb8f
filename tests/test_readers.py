"""Host readers of libkcore against the reference's own readers.

tests/golden/reader_golden.npz holds what the reference's host C++
(peel/shared/graph.cpp + virtual_graph.cpp, peel/src/graph.cpp), compiled
from its sources into oracle/_ref, produced on the quirk files in
tests/golden/readers/ (see make_golden.py).  When oracle/_ref is built the
test also runs it live.  No GPU: only libkcore's host entry points."""
import os
import tempfile

import numpy as np
import pytest

import oracle
import pico_amd as pa

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "reader_golden.npz")
RDIR = os.path.join(HERE, "golden", "readers")
FILES = sorted(f[:-4] for f in os.listdir(RDIR) if f.endswith(".txt")) + ["example"]


def _path(key):
    return os.path.join(HERE, "golden", "example.txt") if key == "example" else \
        os.path.join(RDIR, key + ".txt")


def _coo_layout(key):
    hg = pa.HostGraph.read_coo(_path(key))
    npn, el, parts = hg.vgraph(64)
    return hg, npn, el, parts


@pytest.mark.parametrize("key", FILES)
def test_coo_reader_matches_reference_layout(key):
    z = np.load(GOLD)
    n, e, zero = (int(x) for x in z[key + "__coo_meta"])
    hg, npn, el, parts = _coo_layout(key)
    assert hg.V == n and hg.edges_read == e and hg.has_zero_id == bool(zero)
    np.testing.assert_array_equal(npn, z[key + "__coo_node_pointer"])
    np.testing.assert_array_equal(el, z[key + "__coo_edge_list"])
    np.testing.assert_array_equal(parts.reshape(-1), z[key + "__coo_parts"])


@pytest.mark.parametrize("key", FILES)
def test_dedup_reader_matches_reference(key):
    z = np.load(GOLD)
    if key + "__csr_failed" in z.files:
        # 0 ids wrap to UINT_MAX in src/graph.cpp:89-90 and the reference
        # crashes; libkcore rejects such input with EINVAL instead
        with pytest.raises(pa.KCoreError) as ei:
            pa.HostGraph.read_dedup(_path(key))
        assert ei.value.status == 1
        return
    V, E = (int(x) for x in z[key + "__csr_meta"])
    hg = pa.HostGraph.read_dedup(_path(key))
    assert hg.V == V and hg.nnz == E
    np.testing.assert_array_equal(np.diff(hg.rowptr).astype(np.uint32), z[key + "__csr_degrees"])
    np.testing.assert_array_equal(hg.rowptr.astype(np.uint32), z[key + "__csr_offsets"])
    np.testing.assert_array_equal(hg.col, z[key + "__csr_neighbors"])


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (no reference tree)")
def test_live_reference_readers_on_generated_files():
    # random multigraph edge files, 1-based, with loops and duplicates
    rng = np.random.default_rng(1)
    for trial in range(5):
        n = int(rng.integers(5, 400))
        m = int(rng.integers(1, 3 * n))
        with tempfile.TemporaryDirectory() as wd:
            p = os.path.join(wd, "g%d.txt" % trial)
            with open(p, "w") as f:
                f.write("%%MatrixMarket matrix coordinate integer symmetric\n%d %d %d\n" % (n, n, m))
                for _ in range(m):
                    f.write("%d %d 1\n" % tuple(rng.integers(1, n + 1, 2)))
            coo = oracle.ref_read_coo(p, wd)
            hg = pa.HostGraph.read_coo(p)
            npn, el, parts = hg.vgraph(64)
            assert hg.V == coo["num_nodes"]
            np.testing.assert_array_equal(npn, coo["node_pointer"])
            np.testing.assert_array_equal(el, coo["edge_list"])
            np.testing.assert_array_equal(parts, coo["parts"])
            csr = oracle.ref_read_csr(p, wd)
            hd = pa.HostGraph.read_dedup(p)
            assert hd.V == csr["V"]
            np.testing.assert_array_equal(hd.rowptr.astype(np.uint32), csr["offsets"])
            np.testing.assert_array_equal(hd.col, csr["neighbors"])


def test_writer_round_trip(tmp_path):
    g = pa.HostGraph.er(300, 900, seed=3)
    p = str(tmp_path / "er.txt")
    g.write_edgelist(p)
    with open(p) as f:
        head = [f.readline(), f.readline()]
    assert head[0].startswith("%%MatrixMarket") and head[1].split() == ["300", "300", "900"]
    back = pa.HostGraph.read_dedup(p)
    # trailing isolated ids are dropped by the reference readers (num_nodes =
    # max id); compare over the common range
    V = back.V
    np.testing.assert_array_equal(back.rowptr, g.rowptr[:V + 1])
    np.testing.assert_array_equal(back.col, g.col[:back.nnz])
    coo = pa.HostGraph.read_coo(p)
    assert coo.edges_read == 900


def test_missing_file_is_an_error():
    with pytest.raises(pa.KCoreError) as ei:
        pa.HostGraph.read_coo("/nonexistent/graph.txt")
    assert ei.value.status == 6
