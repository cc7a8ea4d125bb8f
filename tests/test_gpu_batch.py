"""GPU parity of the batched engine (many graphs in one launch) through the C
ABI: every graph's cores equal the oracle's Batagelj-Zaversnik peel bit for
bit, equal the per-graph plain launches, and do not depend on which other
graphs share the launch or on the launch shape.
"""
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


def _slices(hb):
    vo, eo, rp, cl = hb.arrays()
    for g in range(hb.G):
        V = int(vo[g + 1] - vo[g])
        r = rp[vo[g] + g: vo[g] + g + V + 1].astype(np.uint64)
        c = cl[eo[g]: eo[g + 1]]
        yield g, r, c, int(vo[g]), V


def _check_batch(hb, core, kmax, levels_run=None):
    distinct = 0
    for g, r, c, v0, V in _slices(hb):
        want = oracle.bz(r, c)
        np.testing.assert_array_equal(core[v0: v0 + V], want, err_msg="graph %d" % g)
        assert kmax[g] == (int(want.max()) if V else 0), g
        distinct += len(np.unique(want))
    if levels_run is not None:
        # one visited level per distinct core value (empty levels are skipped)
        assert levels_run == distinct


def test_batch_small_mixed(ctx):
    hb = pa.HostBatch.rmat(48, 5, 12, ef_min=1.0, ef_max=24.0, seed=11)
    db = pa.DeviceBatch(ctx, hb)
    core, km, st = db.run()
    assert st["launches"] == 1 and st["graphs"] == 48
    _check_batch(hb, core, km, st["levels_run"])
    db.close()


@pytest.mark.parametrize("block", [256, 512, 1024])
@pytest.mark.parametrize("per_cu", [1, 2])
def test_batch_launch_shapes(ctx, block, per_cu):
    hb = pa.HostBatch.rmat(40, 6, 11, ef_min=2.0, ef_max=32.0, seed=3)
    db = pa.DeviceBatch(ctx, hb)
    core, km, st = db.run(block=block, per_cu=per_cu)
    assert st["block"] == block
    _check_batch(hb, core, km, st["levels_run"])
    db.close()


LDS_SIZES = {256: (4096, 8192), 512: (8192, 16384), 1024: (16384, 32768)}


@pytest.mark.parametrize("block", [256, 512, 1024])
@pytest.mark.parametrize("large", [False, True])
@pytest.mark.parametrize("split", [4, 8])
def test_batch_lds_capacity_and_split(ctx, block, large, split):
    # graphs on both sides of each LDS degree capacity (2^12 .. 2^15
    # vertices) with small ones; every built shape gives the same cores
    gs = [pa.HostGraph.rmat(s, 6, seed=s) for s in (12, 13, 14, 15)]
    gs += [pa.HostGraph.rmat(8, 4, seed=7), pa.HostGraph.er(5000, 20000, seed=2)]
    hb = pa.HostBatch.from_graphs(gs)
    db = pa.DeviceBatch(ctx, hb)
    core, km, st = db.run(block=block, lds_vertices=LDS_SIZES[block][large], split=split)
    assert st["block"] == block
    _check_batch(hb, core, km, st["levels_run"])
    db.close()


def test_batch_shape_validation(ctx):
    hb = pa.HostBatch.rmat(4, 5, 7, seed=1)
    db = pa.DeviceBatch(ctx, hb)
    with pytest.raises(pa.KCoreError):
        db.run(block=512, lds_vertices=32768)  # built for 1024 only
    with pytest.raises(pa.KCoreError):
        db.run(split=6)
    db.close()


def test_batch_graphs_beyond_lds(ctx):
    """Graphs with more vertices than a workgroup keeps in LDS (>16384) take
    the HBM-resident degree path; mixed with small ones."""
    gs = [pa.HostGraph.rmat(15, 8, seed=2), pa.HostGraph.rmat(7, 4, seed=3),
          pa.HostGraph.rmat(16, 4, seed=4), pa.HostGraph.er(20000, 60000, seed=1)]
    hb = pa.HostBatch.from_graphs(gs)
    db = pa.DeviceBatch(ctx, hb)
    for block in (256, 512, 1024):
        core, km, st = db.run(block=block)
        _check_batch(hb, core, km, st["levels_run"])
    db.close()


def test_batch_edge_cases(ctx):
    """Empty graphs, isolated vertices, closed forms and example.txt."""
    gs = []
    for name in graphs.CLOSED_FORMS:
        rp, col, _ = graphs.closed_form(name)
        gs.append(pa.HostGraph.from_csr(rp, col))
    rp, col = graphs.example()
    gs.append(pa.HostGraph.from_csr(rp, col))
    gs.append(pa.HostGraph.from_csr(np.zeros(1, np.uint64), np.zeros(0, np.uint32)))  # V=0
    gs.append(pa.HostGraph.from_csr(np.zeros(6, np.uint64), np.zeros(0, np.uint32)))  # isolated
    hb = pa.HostBatch.from_graphs(gs)
    db = pa.DeviceBatch(ctx, hb)
    core, km, st = db.run()
    _check_batch(hb, core, km, st["levels_run"])
    for name, (g, r, c, v0, V) in zip(graphs.CLOSED_FORMS, _slices(hb)):
        _, _, want = graphs.closed_form(name)
        np.testing.assert_array_equal(core[v0: v0 + V], want, err_msg=name)
    db.close()


def test_batch_alone_equals_in_mix(ctx):
    """A graph's cores are the same decomposed alone and inside a batch."""
    hb = pa.HostBatch.rmat(32, 6, 13, seed=21)
    db = pa.DeviceBatch(ctx, hb)
    core, km, _ = db.run()
    for g in (0, 7, 31):
        one = hb.graph(g)
        hb1 = pa.HostBatch.from_graphs([one])
        d1 = pa.DeviceBatch(ctx, hb1)
        c1, k1, st1 = d1.run()
        vo = hb.vtx_off
        np.testing.assert_array_equal(c1, core[vo[g]: vo[g + 1]])
        assert k1[0] == km[g]
        d1.close()
    db.close()


@pytest.mark.parametrize("method", ["peel_dyn", "peelone_dyn", "gpp", "histocore"])
def test_batch_plain_launches_agree(ctx, method):
    hb = pa.HostBatch.rmat(24, 5, 12, seed=8)
    db = pa.DeviceBatch(ctx, hb)
    core_b, km_b, _ = db.run()
    core_p, km_p, stp = db.run_plain(method)
    assert stp["launches"] == 24
    np.testing.assert_array_equal(core_p, core_b)
    np.testing.assert_array_equal(km_p, km_b)
    _check_batch(hb, core_p, km_p)
    db.close()


def test_batch_autotune_keeps_results(ctx):
    hb = pa.HostBatch.rmat(64, 6, 12, seed=9)
    db = pa.DeviceBatch(ctx, hb)
    best, ms = db.autotune(reps=1)
    assert best["block"] in (256, 512, 1024) and best["per_cu"] >= 1 and ms > 0
    assert best["lds_vertices"] in LDS_SIZES[best["block"]] and best["split"] in (4, 8)
    core, km, st = db.run()
    assert st["block"] == best["block"]
    _check_batch(hb, core, km, st["levels_run"])
    db.close()


def test_batch_upload_arrays_and_validation(ctx):
    hb = pa.HostBatch.rmat(8, 5, 9, seed=4)
    vo, eo, rp, cl = [np.array(a) for a in hb.arrays()]
    db = pa.DeviceBatch(ctx, arrays=(vo, eo, rp, cl))
    core, km, st = db.run()
    _check_batch(hb, core, km, st["levels_run"])
    db.close()
    bad = cl.copy()
    bad[eo[3]] = vo[4] - vo[3]  # local id == V_g
    with pytest.raises(pa.KCoreError, match="graph 3"):
        pa.DeviceBatch(ctx, arrays=(vo, eo, rp, bad))
    rp2 = rp.copy()
    rp2[vo[5] + 5 + 1] = rp2[vo[5] + 5 + 2] + 1  # rows decrease
    with pytest.raises(pa.KCoreError, match="graph 5"):
        pa.DeviceBatch(ctx, arrays=(vo, eo, rp2, cl))


def test_graph_autotune(ctx):
    hg = pa.HostGraph.rmat(14, 16, seed=1)
    dg = ctx.upload(hg)
    want = oracle.bz(hg.rowptr, hg.col)
    for method in ("peelone_dyn", "histocore"):
        best, ms = dg.autotune(method)
        assert best["block"] in (256, 512, 1024) and ms > 0
        core, st = dg.decompose(method)
        assert st["block"] == best["block"]
        np.testing.assert_array_equal(core, want)
    dg.close()


# ------------------------------------------------ device-built graphs
@pytest.mark.parametrize("scale,ef,seed", [(10, 16, 1), (14, 8, 5), (12, 30, 2)])
def test_device_rmat_equals_host(ctx, scale, ef, seed):
    h = pa.HostGraph.rmat(scale, ef, seed=seed)
    dg = ctx.gen_rmat(scale, ef, seed=seed)
    d = dg.download()
    np.testing.assert_array_equal(d.rowptr, h.rowptr)
    np.testing.assert_array_equal(d.col, h.col)
    core, _ = dg.decompose("peelone_dyn")
    np.testing.assert_array_equal(core, oracle.bz(h.rowptr, h.col))
    dg.close()


@pytest.mark.parametrize("width", [4, 8])
def test_device_chunglu_equals_host(ctx, width):
    h = pa.HostGraph.chunglu(1 << 15, beta=2.1, mean_deg=20, cap=4000, seed=4)
    dg = ctx.gen_chunglu(1 << 15, beta=2.1, mean_deg=20, cap=4000, seed=4, rowptr_bytes=width)
    d = dg.download()
    np.testing.assert_array_equal(d.rowptr, h.rowptr)
    np.testing.assert_array_equal(d.col, h.col)
    dg.close()


def test_device_batch_equals_host(ctx):
    hb = pa.HostBatch.rmat(96, 4, 12, ef_min=1.0, ef_max=32.0, seed=17)
    db = ctx.gen_rmat_batch(96, 4, 12, ef_min=1.0, ef_max=32.0, seed=17)
    d = db.download()
    for x, y in zip(d.arrays(), hb.arrays()):
        np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(d.scales, hb.scales)
    core, km, st = db.run()
    _check_batch(hb, core, km, st["levels_run"])
    db.close()
