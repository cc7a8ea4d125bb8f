"""k-core subgraph extraction and degree statistics on resident graphs
(the consumers main.cu:10-63 feeds: PrintMax, STDdegrees, the kmax-core),
checked against numpy restatements on the same graphs."""
import os

import numpy as np
import pytest

import oracle
import pico_amd as pa
from tests import graphs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = pa.Context(0)
    yield c
    c.close()


def _host_subgraph(rp, col, core, k):
    rp = np.asarray(rp, np.int64)
    col = np.asarray(col, np.int64)
    keep = core >= k
    ids = np.flatnonzero(keep).astype(np.uint32)
    new = np.cumsum(keep) - 1
    rows, cols = [0], []
    for v in ids:
        nb = col[rp[v]:rp[v + 1]]
        nb = nb[keep[nb]]  # row order kept
        cols.extend(new[nb].tolist())
        rows.append(len(cols))
    return ids, np.array(rows, np.uint64), np.array(cols, np.uint32)


@pytest.mark.parametrize("gen", ["example", "rmat12", "chunglu_wide", "petersen"])
def test_kcore_subgraph_matches_numpy(ctx, gen):
    if gen == "example":
        rp, col = graphs.example()
        hg = pa.HostGraph.from_csr(rp, col)
        dg = ctx.upload(hg)
    elif gen == "rmat12":
        hg = pa.HostGraph.rmat(12, 16, seed=11)
        dg = ctx.upload(hg)
    elif gen == "petersen":
        rp, col, _ = graphs.closed_form("petersen")
        hg = pa.HostGraph.from_csr(rp, col)
        dg = ctx.upload(hg)
    else:
        hg = pa.HostGraph.chunglu(1 << 13, beta=2.1, mean_deg=20.0, cap=2000.0, seed=2)
        dg = ctx.upload(hg, rowptr_bytes=8)
    core, st = dg.decompose("peelone_dyn")
    want_core = oracle.bz(hg.rowptr, hg.col)
    np.testing.assert_array_equal(core, want_core)
    kmax = int(want_core.max()) if hg.V else 0
    for k in sorted({0, 1, 2, max(kmax // 2, 1), kmax, kmax + 1}):
        sub, ids = dg.kcore_subgraph(k)
        wid, wrp, wcol = _host_subgraph(hg.rowptr, hg.col, want_core, k)
        np.testing.assert_array_equal(ids, wid)
        h = sub.download()
        np.testing.assert_array_equal(h.rowptr, wrp)
        np.testing.assert_array_equal(h.col, wcol)
        if len(ids) and k >= 1:
            # the k-core's own core numbers are the graph's
            sc, _ = sub.decompose("histocore")
            np.testing.assert_array_equal(sc, want_core[ids])
        sub.close()
    dg.close()


def test_subgraph_needs_a_decomposition(ctx):
    dg = ctx.upload(pa.HostGraph.rmat(8, 8, seed=1))
    with pytest.raises(pa.KCoreError):
        dg.kcore_subgraph(1)
    dg.close()


def test_subgraph_after_edge_updates(ctx):
    hg = pa.HostGraph.rmat(10, 8, seed=4)
    dg = ctx.upload(hg)
    dg.decompose("peelone_dyn")
    rng = np.random.default_rng(5)
    u = rng.integers(0, hg.V, 200).astype(np.uint32)
    v = rng.integers(0, hg.V, 200).astype(np.uint32)
    m = u != v
    core, _ = dg.update("insert", u[m], v[m])
    back = dg.download()
    want_core = oracle.bz(back.rowptr, back.col)
    np.testing.assert_array_equal(core, want_core)
    k = int(want_core.max())
    sub, ids = dg.kcore_subgraph(k)
    wid, wrp, wcol = _host_subgraph(back.rowptr, back.col, want_core, k)
    np.testing.assert_array_equal(ids, wid)
    h = sub.download()
    np.testing.assert_array_equal(h.rowptr, wrp)
    np.testing.assert_array_equal(h.col, wcol)
    sub.close()
    dg.close()


@pytest.mark.parametrize("gen", ["example", "rmat12", "chunglu_wide", "empty"])
def test_degree_stats(ctx, gen):
    if gen == "example":
        hg = pa.HostGraph.read_dedup(os.path.join(graphs.GOLDEN, "example.txt"))
        dg = ctx.upload(hg)
    elif gen == "rmat12":
        hg = pa.HostGraph.rmat(12, 16, seed=11)
        dg = ctx.upload(hg)
    elif gen == "empty":
        hg = pa.HostGraph.from_csr([0, 0, 0], [])
        dg = ctx.upload(hg)
    else:
        hg = pa.HostGraph.chunglu(1 << 13, beta=2.1, mean_deg=20.0, cap=2000.0, seed=2)
        dg = ctx.upload(hg, rowptr_bytes=8)
    ds = dg.degree_stats()
    deg = np.diff(np.asarray(hg.rowptr, np.int64)).astype(np.float64)
    assert ds["V"] == hg.V and ds["nnz"] == hg.nnz
    assert ds["dmax"] == (int(deg.max()) if hg.V else 0)
    assert ds["mean"] == pytest.approx(deg.mean() if hg.V else 0.0, rel=1e-12, abs=1e-12)
    # population standard deviation, as STDdegrees computes it
    assert ds["stddev"] == pytest.approx(deg.std() if hg.V else 0.0, rel=1e-9, abs=1e-12)
    if gen == "example":
        assert ds["dmax"] == 3 and ds["stddev"] == pytest.approx(0.4)
    dg.close()
