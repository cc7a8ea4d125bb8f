"""libkcore without a GPU: the library loads, exports every KC_API symbol of
include/kcore.h, and its host entry points (generators, CSR builder, layout
export) behave.  No compute call touches a device here."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import oracle
import pico_amd as pa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = []
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            src = open(os.path.join(ROOT, "include", h)).read()
            names += re.findall(r"KC_API[^;(]*?\b(kc_\w+)\s*\(", src, flags=re.S)
    return sorted(set(names))


def test_every_declared_symbol_is_exported():
    names = _declared()
    assert len(names) > 20
    L = ctypes.CDLL(pa.kcore.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_no_undeclared_exports():
    out = subprocess.run(["nm", "-D", "--defined-only", pa.kcore.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T kc_" in l}
    assert exported <= set(_declared()), exported - set(_declared())


def _is_symmetric_simple(g):
    rp, col = g.rowptr, g.col
    V = g.V
    src = np.repeat(np.arange(V, dtype=np.uint64), np.diff(rp).astype(np.int64))
    assert not np.any(src == col), "self loop"
    for v in range(0, V, max(1, V // 200)):
        row = col[rp[v]:rp[v + 1]]
        assert np.all(np.diff(row.astype(np.int64)) > 0), "unsorted or duplicate row"
    fwd = set(zip(src.tolist()[:20000], col.tolist()[:20000]))
    key = (src.astype(np.uint64) << np.uint64(32)) | col.astype(np.uint64)
    rev = (col.astype(np.uint64) << np.uint64(32)) | src.astype(np.uint64)
    assert np.array_equal(np.sort(key), np.sort(rev))
    return fwd


def test_er_generator_exact_and_simple():
    g = pa.HostGraph.er(10000, 50000, seed=0)
    assert g.V == 10000 and g.nnz == 100000
    _is_symmetric_simple(g)


def test_rmat_generator_properties():
    g = pa.HostGraph.rmat(12, 16, seed=1)
    assert g.V == 4096 and 0 < g.nnz <= 2 * 16 * 4096
    _is_symmetric_simple(g)
    # the a=0.57 quadrant concentrates degree on low ids
    d = np.diff(g.rowptr)
    assert d[:64].mean() > 4 * d.mean()


def test_generators_independent_of_thread_count():
    code = ("import sys; sys.path.insert(0, %r); import pico_amd as pa, numpy as np;"
            "g = pa.HostGraph.rmat(13, 8, seed=5); c = pa.HostGraph.chunglu(5000, seed=2);"
            "print(int(np.uint64(np.sum(g.col.astype(np.uint64) * np.arange(g.nnz, dtype=np.uint64)))),"
            " g.nnz, c.nnz, int(c.col.astype(np.uint64).sum()))" % ROOT)
    outs = []
    for t in ("1", "7"):
        env = dict(os.environ, KC_HOST_THREADS=t)
        outs.append(subprocess.run([sys.executable, "-c", code], env=env, capture_output=True,
                                   text=True, check=True).stdout)
    assert outs[0] == outs[1]


def test_chunglu_hub_degrees():
    g = pa.HostGraph.chunglu(1 << 14, beta=2.1, mean_deg=20.0, cap=3000.0, seed=9)
    d = np.diff(g.rowptr)
    _is_symmetric_simple(g)
    assert d.max() <= 2 * 3000 and d.max() > 1000
    assert 10 < d.mean() < 21


def test_from_pairs_symmetrizes_and_dedups():
    src = np.array([0, 1, 1, 2, 3, 3], np.uint32)
    dst = np.array([1, 0, 2, 2, 0, 0], np.uint32)
    g = pa.HostGraph.from_pairs(4, src, dst)
    assert g.rowptr.tolist() == [0, 2, 4, 5, 6]
    assert g.col.tolist() == [1, 3, 0, 2, 1, 0]


def test_from_pairs_rejects_out_of_range():
    with pytest.raises(pa.KCoreError):
        pa.HostGraph.from_pairs(3, np.array([0], np.uint32), np.array([5], np.uint32))


def test_from_csr_validates():
    with pytest.raises(pa.KCoreError):
        pa.HostGraph.from_csr(np.array([0, 1, 3], np.uint64), np.array([1, 0, 9], np.uint32))
    with pytest.raises(pa.KCoreError):
        pa.HostGraph.from_csr(np.array([0, 2, 1], np.uint64), np.array([1, 0], np.uint32))


def test_vgraph_layout_matches_definition():
    g = pa.HostGraph.er(500, 3000, seed=4)
    npn, el, parts = g.vgraph(64)
    rp, col = g.rowptr, g.col
    assert np.array_equal(npn, (rp[:-1] + np.arange(g.V)).astype(np.uint32))
    for v in range(0, g.V, 37):
        b = int(npn[v])
        assert el[b] == rp[v + 1] - rp[v]
        assert np.array_equal(el[b + 1:b + 1 + el[b]], col[rp[v]:rp[v + 1]])
    d = np.diff(rp)
    assert len(parts) == int(((d + 63) // 64).sum())


def test_write_cores_format(tmp_path):
    p = str(tmp_path / "cores.txt")
    core = np.array([2, 0, 5], np.uint32)
    r = pa.lib().kc_write_cores(os.fsencode(p), core.ctypes.data_as(ctypes.c_void_p), 3)
    assert r == 0
    assert open(p).read() == "0 2\n1 0\n2 5\n"


def test_device_entry_points_fail_cleanly_without_gpu():
    # no GPU in this container: context creation reports an error, never exits
    try:
        ctx = pa.Context(0)
    except pa.KCoreError as ex:
        assert ex.status == 3
    else:
        ctx.close()
