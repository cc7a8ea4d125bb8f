"""GPU parity of the h-index family (HistoCore, CntCore, NbrCore) against the
oracle, through the C ABI: core numbers bit-exact vs Batagelj-Zaversnik, the
"Number of iterations" equal to the oracle's restatement of each reference
host loop (HistoCore.cu:301-341, CntCore.cu:344-417, NbrCore.cu:244-297) and
to the plain Jacobi round count, and HistoCore's final histogram equal word
for word to the restatement's.
"""
import numpy as np
import pytest

import oracle
import pico_amd as pa
from pico_amd.kcore import device_histogram
from tests import graphs

pytestmark = pytest.mark.gpu

HIDX = {"histocore": oracle.histocore, "cntcore": oracle.cntcore, "nbrcore": oracle.nbrcore}


@pytest.fixture(scope="module")
def ctx():
    c = pa.Context(0)
    yield c
    c.close()


def _check(ctx, rowptr, col, method, width=4, hist=False, **opts):
    hg = pa.HostGraph.from_csr(rowptr, col)
    dg = ctx.upload(hg, rowptr_bytes=width)
    try:
        core, st = dg.decompose(method, **opts)
        if hist:
            dev_hist = device_histogram(dg, hg.nnz + hg.V)
    finally:
        dg.close()
    want = oracle.bz(rowptr, col)
    np.testing.assert_array_equal(core, want)
    if method == "histocore":
        ref_core, it, ref_hist = oracle.histocore(rowptr, col, want_hist=True)
        if hist:
            np.testing.assert_array_equal(dev_hist, ref_hist)
    else:
        ref_core, it = HIDX[method](rowptr, col)
    _, jac = oracle.hindex_jacobi(rowptr, col)
    assert st["iterations"] == it == jac, (st["iterations"], it, jac)
    return core, st


METHODS = ["histocore", "cntcore", "nbrcore"]


@pytest.mark.parametrize("method", METHODS)
def test_example_txt(ctx, method):
    rp, col = graphs.example()
    core, _ = _check(ctx, rp, col, method, hist=True)
    assert core.tolist() == [2, 2, 2, 2, 2]


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("name", graphs.CLOSED_FORMS)
def test_closed_forms(ctx, method, name):
    rp, col, expect = graphs.closed_form(name)
    core, _ = _check(ctx, rp, col, method)
    np.testing.assert_array_equal(core, expect)


@pytest.mark.parametrize("method", METHODS)
def test_er_config(ctx, method):
    hg = pa.HostGraph.er(10000, 50000, seed=0)
    _check(ctx, hg.rowptr, hg.col, method, hist=True)


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("scale", [10, 14])
def test_rmat_small(ctx, method, scale):
    hg = pa.HostGraph.rmat(scale, 16, seed=scale)
    _check(ctx, hg.rowptr, hg.col, method, hist=(method == "histocore"))


@pytest.mark.parametrize("method", METHODS)
def test_chunglu_hubs_wide_rowptr(ctx, method):
    # hubs above 4096 neighbours exercise the workgroup-level paths
    hg = pa.HostGraph.chunglu(1 << 15, beta=2.1, mean_deg=20.0, cap=20000.0, seed=5)
    _check(ctx, hg.rowptr, hg.col, method, width=8, hist=(method == "histocore"))


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("block", [256, 512, 1024])
def test_launch_shapes(ctx, method, block):
    hg = pa.HostGraph.rmat(12, 16, seed=7)
    _check(ctx, hg.rowptr, hg.col, method, block=block)
    _check(ctx, hg.rowptr, hg.col, method, block=block, grid=3)


@pytest.mark.parametrize("method", METHODS)
def test_multigraph_from_coo_reader(ctx, method, tmp_path):
    # the COO reader keeps duplicate edges (shared/graph.cpp:49-81): cores of
    # the multigraph, as every reference method computes them
    p = tmp_path / "dup.txt"
    p.write_text("%%header\n5 5 9\n1 2 1\n2 1 1\n1 3 1\n2 3 1\n3 4 1\n4 5 1\n5 3 1\n4 5 1\n2 2 1\n")
    hg = pa.HostGraph.read_coo(str(p))
    assert hg.multigraph and hg.edges_read == 8
    _check(ctx, hg.rowptr, hg.col, method)


@pytest.mark.parametrize("flags", [0, pa.kcore.KC_OPT_NO_DEGSORT])
def test_histocore_adjacency_order_does_not_matter(ctx, flags):
    # the degree-ordered adjacency only prunes update work; cores, rounds and
    # the final histogram are the same with and without it
    hg = pa.HostGraph.chunglu(1 << 14, beta=2.1, mean_deg=20.0, cap=6000.0, seed=8)
    _check(ctx, hg.rowptr, hg.col, "histocore", hist=True, flags=flags)


@pytest.mark.parametrize("method", METHODS)
def test_atomic_count_bounds(ctx, method):
    hg = pa.HostGraph.rmat(13, 16, seed=5)
    dg = ctx.upload(hg)
    try:
        _, st = dg.decompose(method)
    finally:
        dg.close()
    assert st["changed"] > 0 and st["atomics"] > 0
    if method == "histocore":
        # initial histograms: at most one atomic per adjacency entry; each
        # update-walk entry at most a bin decrement and a bin increment, or
        # their summed forms
        assert st["atomics"] <= hg.nnz + 2 * st["edges_walked"]


def _clique_with_leaves(n, leaves):
    """K_n whose members carry `leaves` pendant vertices each: every clique
    member starts at degree n-1+leaves and drops to n-1 in the first round."""
    iu, ju = np.triu_indices(n, 1)
    owner = np.repeat(np.arange(n), leaves)
    leaf = n + np.arange(n * leaves)
    a = np.concatenate([iu, owner]).astype(np.int64)
    b = np.concatenate([ju, leaf]).astype(np.int64)
    src, dst = np.concatenate([a, b]), np.concatenate([b, a])
    order = np.lexsort((dst, src))
    V = n + n * leaves
    rowptr = np.zeros(V + 1, np.uint64)
    rowptr[1:] = np.cumsum(np.bincount(src, minlength=V))
    return rowptr, dst[order].astype(np.uint32)


@pytest.mark.parametrize("n,block,grid", [(1100, 512, 0), (1100, 512, 3), (4200, 256, 0)])
def test_histocore_hub_rebuild_paths(ctx, n, block, grid):
    # a round that reaches many hub edges lets every hub (more than 1,024
    # neighbours) pull its changed neighbours' updates instead of receiving
    # patches: cores of 1,099 (above a wave's 1,024 LDS bins) go to the
    # workgroup, a workgroup's overflow of such hubs (grid 3) to one wave with
    # HBM atomics for the bins from 1,024 up, and 4199 (above a 256-thread
    # group's 4,096 LDS bins) to HBM atomics; bin 0 and the bins above the
    # core must come out as the reference leaves them
    rp, col = _clique_with_leaves(n, 10)
    _, st = _check(ctx, rp, col, "histocore", hist=True, block=block, grid=grid)
    assert st["rebuild_rounds"] > 0, st


def _union(*parts):
    """Disjoint union of CSR graphs (vertex ids shifted part by part)."""
    rps, cols, off = [np.zeros(1, np.uint64)], [], 0
    for rp, col in parts:
        rps.append(rp[1:] + rps[-1][-1])
        cols.append(col.astype(np.uint32) + np.uint32(off))
        off += len(rp) - 1
    return np.concatenate(rps), np.concatenate(cols)


@pytest.mark.parametrize("block", [256, 512, 1024])
def test_histocore_wave_rebuild_and_hub_aggregation(ctx, block):
    # K_100 whose members carry 1,100 leaves each: hubs of 1,199 neighbours
    # (above the 1,024 of a pull hub, at most 4,096) with core 99 (below a
    # wave's 1,024 LDS bins), pulled by one wave each when the first round's changed
    # hubs reach all of the hub adjacency; beside it K_300 with 50 leaves
    # each: 349 neighbours (257..512), whose bin updates are summed per
    # workgroup in LDS instead of patched one atomic at a time
    rp, col = _union(_clique_with_leaves(100, 1100), _clique_with_leaves(300, 50))
    deg = np.diff(rp)
    assert ((deg > 1024) & (deg <= 4096)).sum() == 100
    assert ((deg > 256) & (deg <= 512)).sum() == 300
    core, st = _check(ctx, rp, col, "histocore", hist=True, block=block)
    assert st["rebuild_rounds"] > 0, st
    assert int(core[:100].max()) == 99 and int(core[110100:110400].min()) == 299


@pytest.mark.parametrize("pull_div", [1, 0, 12, 0xFFFFFFFF])
@pytest.mark.parametrize("graph", ["rmat", "chunglu"])
def test_histocore_pull_rounds_match_patches(ctx, pull_div, graph):
    # hubs apply their changed neighbours' updates by a pull over their own
    # row in some rounds (all of them with 0xffffffff, about none with 1); the
    # cores, the rounds and every word of the final histogram are the same
    # whichever rounds pull
    if graph == "rmat":
        hg = pa.HostGraph.rmat(15, 16, seed=11)
    else:
        hg = pa.HostGraph.chunglu(1 << 15, beta=2.1, mean_deg=20.0, cap=12000.0, seed=9)
    _, st = _check(ctx, hg.rowptr, hg.col, "histocore", hist=True, pull_div=pull_div)
    if pull_div == 0xFFFFFFFF:
        assert st["rebuild_rounds"] > 0, st


@pytest.mark.parametrize("leaves", [255, 256, 65535, 65536])
def test_degree_sort_byte_boundaries(ctx, leaves):
    # the degree-ordered rows sort only the bytes the largest degree needs: a
    # hub of 255 / 256 / 65,535 / 65,536 neighbours (one, two or three key
    # bytes) beside a clique whose members carry leaves; the rows must come
    # out ordered either way
    rp, col = _union(_clique_with_leaves(1, leaves), _clique_with_leaves(40, 3),
                     _clique_with_leaves(12, 200))
    for method in ("histocore", "nbrcore"):
        _check(ctx, rp, col, method, hist=(method == "histocore"))
