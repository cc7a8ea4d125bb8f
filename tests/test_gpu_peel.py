"""GPU parity of the peel family (gpp, peelone, peel_dyn, peelone_dyn) against
the oracle, through the C ABI.  Core numbers are compared bit-exactly; GPP's
iteration count must equal the reference loop's (oracle restatement of
gppeel.cu:186-228); the bucket peels report kmax+1 levels (peel-dyn.cc:155-176).
"""
import os

import numpy as np
import pytest

import oracle
import pico_amd as pa
from tests import graphs

pytestmark = pytest.mark.gpu

PEELS = ["gpp", "peelone", "peel_dyn", "peelone_dyn"]


@pytest.fixture(scope="module")
def ctx():
    c = pa.Context(0)
    yield c
    c.close()


def _check(ctx, rowptr, col, method, width=4, **opts):
    hg = pa.HostGraph.from_csr(rowptr, col)
    dg = ctx.upload(hg, rowptr_bytes=width)
    try:
        core, st = dg.decompose(method, **opts)
    finally:
        dg.close()
    want = oracle.bz(rowptr, col)
    if method == "gpp":
        ref, it = oracle.gpp(rowptr, col, compat=bool(opts.get("gpp_compat")))
        assert st["iterations"] == it, (st["iterations"], it)
        np.testing.assert_array_equal(core, ref)
        if not opts.get("gpp_compat"):
            np.testing.assert_array_equal(core, want)
    else:
        np.testing.assert_array_equal(core, want)
    if len(rowptr) > 1:
        assert st["kmax"] == int(want.max()) if method != "gpp" or not opts.get("gpp_compat") \
            else True
    if method in ("peel_dyn", "peelone_dyn"):
        _, levels = oracle.bucket_peel(rowptr, col)
        assert st["iterations"] == levels
    return core, st


@pytest.mark.parametrize("method", PEELS)
def test_example_txt(ctx, method):
    rp, col = graphs.example()
    core, _ = _check(ctx, rp, col, method)
    assert core.tolist() == [2, 2, 2, 2, 2]


@pytest.mark.parametrize("method", PEELS)
@pytest.mark.parametrize("name", ["k1", "k2", "k7", "cycle9", "star12", "petersen", "empty5",
                                  "forest", "k33", "two_cliques"])
def test_closed_forms(ctx, method, name):
    rp, col, expect = graphs.closed_form(name)
    core, _ = _check(ctx, rp, col, method)
    np.testing.assert_array_equal(core, expect)


@pytest.mark.parametrize("method", PEELS)
def test_er_config(ctx, method):
    hg = pa.HostGraph.er(10000, 50000, seed=0)
    _check(ctx, hg.rowptr, hg.col, method)


@pytest.mark.parametrize("method", PEELS)
@pytest.mark.parametrize("scale", [10, 14])
def test_rmat_small(ctx, method, scale):
    hg = pa.HostGraph.rmat(scale, 16, seed=scale)
    _check(ctx, hg.rowptr, hg.col, method)


@pytest.mark.parametrize("method", PEELS)
def test_chunglu_small_wide_rowptr(ctx, method):
    hg = pa.HostGraph.chunglu(1 << 14, beta=2.1, mean_deg=20.0, cap=4000.0, seed=5)
    _check(ctx, hg.rowptr, hg.col, method, width=8)


@pytest.mark.parametrize("method", ["peel_dyn", "peelone", "gpp"])
@pytest.mark.parametrize("block", [256, 512, 1024])
def test_launch_shapes(ctx, method, block):
    hg = pa.HostGraph.rmat(12, 16, seed=7)
    _check(ctx, hg.rowptr, hg.col, method, block=block)


@pytest.mark.parametrize("method", ["peelone_dyn", "peelone", "gpp"])
def test_hub_chunks_and_spill(ctx, method):
    # hub_split=64 cuts every vertex of degree > 64 into virtual nodes; one
    # workgroup overflows its LDS frontier and spills to the shared list.
    hg = pa.HostGraph.rmat(13, 16, seed=9)
    _check(ctx, hg.rowptr, hg.col, method, hub_split=64)
    _check(ctx, hg.rowptr, hg.col, method, grid=1, block=256)
    _check(ctx, hg.rowptr, hg.col, method, grid=3, hub_split=64)


def test_gpp_compat_isolated(ctx):
    rp, col, _ = graphs.closed_form("empty5")
    core, _ = _check(ctx, rp, col, "gpp", gpp_compat=1)
    assert core.tolist() == [1, 1, 1, 1, 1]


def test_queue_overflow_reports_error(ctx):
    hg = pa.HostGraph.rmat(12, 16, seed=11)
    dg = ctx.upload(hg)
    with pytest.raises(pa.KCoreError) as ei:
        dg.decompose("peelone_dyn", grid=1, block=256, queue_cap=8, hub_split=64)
    assert ei.value.status == 4
    # the graph stays usable after the error
    core, _ = dg.decompose("peelone_dyn")
    np.testing.assert_array_equal(core, oracle.bz(hg.rowptr, hg.col))
    dg.close()


def test_ref_peel_entry_points(ctx):
    hg = pa.HostGraph.er(2000, 9000, seed=4)
    want = oracle.bz(hg.rowptr, hg.col)
    for which in ("peel_dyn", "peelone_dyn"):
        ms, kmax, core = pa.ref_peel(hg, which)
        assert ms >= 0 and kmax == int(want.max())
        np.testing.assert_array_equal(core, want)


def test_vgraph_layout_entry(ctx):
    hg = pa.HostGraph.rmat(11, 8, seed=2)
    npn, el, _ = hg.vgraph(64)
    core, _ = pa.vgraph_decompose(npn, el, "peelone_dyn")
    np.testing.assert_array_equal(core, oracle.bz(hg.rowptr, hg.col))


def test_repeat_is_deterministic(ctx):
    hg = pa.HostGraph.rmat(15, 16, seed=21)
    dg = ctx.upload(hg)
    first = None
    for _ in range(3):
        core, st = dg.decompose("peelone_dyn")
        if first is None:
            first = core
        np.testing.assert_array_equal(core, first)
    dg.close()


@pytest.mark.parametrize("method", PEELS)
def test_atomic_count_bounds(ctx, method):
    # every walked adjacency entry issues at most a decrement plus a restore
    # or a claim; every vertex at most one scan claim
    hg = pa.HostGraph.rmat(13, 16, seed=5)
    dg = ctx.upload(hg)
    try:
        core, st = dg.decompose(method, flags=pa.kcore.KC_OPT_COUNT)
        _, plain = dg.decompose(method)
    finally:
        dg.close()
    np.testing.assert_array_equal(core, oracle.bz(hg.rowptr, hg.col))
    assert st["edges_walked"] > 0
    assert 0 < st["atomics"] <= 3 * st["edges_walked"] + hg.V
    assert plain["atomics"] == 0  # the default variant does not count


@pytest.mark.parametrize("method", ["peel_dyn", "peelone_dyn"])
@pytest.mark.parametrize("div", [2, 4, 16])
@pytest.mark.parametrize("gen", ["rmat15", "chunglu_wide"])
def test_compacted_stages(ctx, method, div, gen):
    # the bucket peels continue on compacted live subgraphs: same cores, same
    # visited levels as without compaction, fewer adjacency entries walked
    if gen == "rmat15":
        hg = pa.HostGraph.rmat(15, 16, seed=9)
        dg = ctx.upload(hg)
    else:
        hg = pa.HostGraph.chunglu(1 << 15, beta=2.1, mean_deg=20.0, cap=4000.0, seed=4)
        dg = ctx.upload(hg, rowptr_bytes=8)
    try:
        core, st = dg.decompose(method, compact_div=div)
        ref, st0 = dg.decompose(method, compact_div=0xFFFFFFFF)
        for block in (256, 1024):
            again, _ = dg.decompose(method, compact_div=div, block=block)
            np.testing.assert_array_equal(again, core)
        for sg in (1, 8, 1000):  # compacted stages on 1, 8 or all workgroups
            again, _ = dg.decompose(method, compact_div=div, stage_grid=sg)
            np.testing.assert_array_equal(again, core)
    finally:
        dg.close()
    want = oracle.bz(hg.rowptr, hg.col)
    np.testing.assert_array_equal(core, want)
    np.testing.assert_array_equal(ref, want)
    assert st["iterations"] == st0["iterations"] == int(want.max()) + 1
    assert st["levels_run"] == st0["levels_run"]
    assert st["edges_walked"] <= st0["edges_walked"]
    assert st0["stages"] == 1
    if div <= 4:  # V/div >= 4,096 live vertices: at least one compaction
        assert st["stages"] >= 2


@pytest.mark.parametrize("gen", ["rmat15", "chunglu15"])
def test_peelone_compacted_stages(ctx, gen):
    # PeelOne continues on compacted live subgraphs too: same cores, kmax and
    # "Number of iterations" as in one launch (each stopped launch's count of
    # the loop-ending sweep is taken back), at least one compaction
    if gen == "rmat15":
        hg = pa.HostGraph.rmat(15, 16, seed=9)
        dg = ctx.upload(hg)
    else:
        hg = pa.HostGraph.chunglu(1 << 15, beta=2.1, mean_deg=20.0, cap=4000.0, seed=4)
        dg = ctx.upload(hg, rowptr_bytes=8)
    try:
        core, st = dg.decompose("peelone", compact_div=4)
        ref, st0 = dg.decompose("peelone", compact_div=0xFFFFFFFF)
    finally:
        dg.close()
    want = oracle.bz(hg.rowptr, hg.col)
    np.testing.assert_array_equal(core, want)
    np.testing.assert_array_equal(ref, want)
    assert st["kmax"] == st0["kmax"] == int(want.max())
    assert st0["stages"] == 1 and st["stages"] >= 2
    assert st["iterations"] == st0["iterations"]
    assert st["levels_run"] == st0["levels_run"]
