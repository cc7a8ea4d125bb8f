"""Binary CSR files and the content-keyed reader cache (the replacement for
the text ./output/serialized-<name> files of peel/src/graph.cpp:9-54): a
round trip is exact, entries are found by content and reader (not by file
name), and a damaged or stale entry is re-parsed instead of served.  CPU
only."""
import os
import shutil

import numpy as np
import pytest

import pico_amd as pa

EXAMPLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "example.txt")


def _same(a, b):
    for k in ("V", "nnz", "edges_read", "has_zero_id", "multigraph"):
        assert getattr(a, k) == getattr(b, k), k
    np.testing.assert_array_equal(a.rowptr, b.rowptr)
    np.testing.assert_array_equal(a.col, b.col)


def _write_rmat(path, scale=9, seed=3):
    g = pa.HostGraph.rmat(scale, 8, seed=seed)
    g.write_edgelist(str(path))
    return g


def test_save_load_round_trip(tmp_path):
    for g in (pa.HostGraph.read_coo(EXAMPLE), pa.HostGraph.read_dedup(EXAMPLE),
              pa.HostGraph.rmat(10, 8, seed=2), pa.HostGraph.from_csr([0], [])):
        p = tmp_path / "g.bin"
        g.save(str(p))
        _same(g, pa.HostGraph.load(str(p)))


def test_cached_read_equals_reader_and_hits_by_content(tmp_path):
    src = tmp_path / "a.txt"
    _write_rmat(src)
    cache = tmp_path / "cache"
    cache.mkdir()
    for reader, direct in (("coo", pa.HostGraph.read_coo), ("dedup", pa.HostGraph.read_dedup)):
        want = direct(str(src))
        g1, hit1 = pa.HostGraph.read_cached(str(src), reader, str(cache))
        g2, hit2 = pa.HostGraph.read_cached(str(src), reader, str(cache))
        assert (hit1, hit2) == (False, True)
        _same(g1, want)
        _same(g2, want)
    # one entry per (content, reader)
    assert len(os.listdir(cache)) == 2
    # same content under another name: a hit; the reference's name key would miss
    other = tmp_path / "b.txt"
    shutil.copy(src, other)
    _, hit = pa.HostGraph.read_cached(str(other), "dedup", str(cache))
    assert hit


def test_changed_file_of_same_name_misses(tmp_path):
    src = tmp_path / "g.txt"
    cache = tmp_path / "cache"
    cache.mkdir()
    _write_rmat(src, seed=3)
    g1, _ = pa.HostGraph.read_cached(str(src), "coo", str(cache))
    _write_rmat(src, seed=4)  # same name, new content
    g2, hit = pa.HostGraph.read_cached(str(src), "coo", str(cache))
    assert not hit
    _same(g2, pa.HostGraph.read_coo(str(src)))
    assert g1.nnz != g2.nnz or not np.array_equal(g1.col, g2.col)


@pytest.mark.parametrize("damage", ["flip", "truncate", "garbage"])
def test_damaged_entry_is_reparsed(tmp_path, damage):
    src = tmp_path / "g.txt"
    _write_rmat(src)
    cache = tmp_path / "cache"
    cache.mkdir()
    pa.HostGraph.read_cached(str(src), "dedup", str(cache))
    (entry,) = [cache / f for f in os.listdir(cache)]
    data = bytearray(entry.read_bytes())
    if damage == "flip":
        data[-5] ^= 0x40  # a col entry: only the payload hash catches it
    elif damage == "truncate":
        data = data[: len(data) - 12]
    else:
        data[:8] = b"notacsr!"
    entry.write_bytes(bytes(data))
    with pytest.raises(pa.KCoreError):
        pa.HostGraph.load(str(entry))
    g, hit = pa.HostGraph.read_cached(str(src), "dedup", str(cache))
    assert not hit
    _same(g, pa.HostGraph.read_dedup(str(src)))
    # the entry was rewritten and is good again
    _, hit = pa.HostGraph.read_cached(str(src), "dedup", str(cache))
    assert hit


def test_errors(tmp_path):
    with pytest.raises(pa.KCoreError):
        pa.HostGraph.read_cached(str(tmp_path / "missing.txt"), "coo", str(tmp_path))
    with pytest.raises(pa.KCoreError):
        pa.HostGraph.load(str(tmp_path / "missing.bin"))
    src = tmp_path / "g.txt"
    _write_rmat(src)
    with pytest.raises(pa.KCoreError):  # cache directory that does not exist
        pa.HostGraph.read_cached(str(src), "coo", str(tmp_path / "no" / "such" / "dir"))


def test_cache_keeps_reader_quirks(tmp_path):
    # the COO reader keeps duplicate lines and shifts 1-based ids only when no
    # id is 0; the entry must carry those flags through
    src = tmp_path / "dup.txt"
    src.write_text("% header\n% header\n1 2 1\n2 1 1\n2 3 1\n3 1 1\n")
    cache = tmp_path / "cache"
    cache.mkdir()
    want = pa.HostGraph.read_coo(str(src))
    pa.HostGraph.read_cached(str(src), "coo", str(cache))
    g, hit = pa.HostGraph.read_cached(str(src), "coo", str(cache))
    assert hit
    _same(g, want)
    assert g.multigraph and want.multigraph
