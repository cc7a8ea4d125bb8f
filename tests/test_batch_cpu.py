"""Host side of the batched path (no GPU): the R-MAT batch generator's draws
and determinism, the concatenated layout, and graph extraction."""
import numpy as np

import oracle
import pico_amd as pa


def test_batch_layout_and_determinism():
    a = pa.HostBatch.rmat(40, 5, 11, ef_min=4.0, ef_max=32.0, seed=7)
    b = pa.HostBatch.rmat(40, 5, 11, ef_min=4.0, ef_max=32.0, seed=7)
    for x, y in zip(a.arrays(), b.arrays()):
        np.testing.assert_array_equal(x, y)
    vo, eo, rp, cl = a.arrays()
    sc, mg = a.scales, a.m_gen
    assert sc.min() >= 5 and sc.max() <= 11
    np.testing.assert_array_equal(np.diff(vo), 1 << sc.astype(np.uint64))
    for g in range(a.G):
        ef = mg[g] / float(1 << int(sc[g]))
        assert 4.0 - 1e-9 <= ef <= 32.0 + 1e-9
        V = int(vo[g + 1] - vo[g])
        r = rp[vo[g] + g: vo[g] + g + V + 1]
        assert r[0] == 0 and r[-1] == eo[g + 1] - eo[g]
        assert np.all(np.diff(r.astype(np.int64)) >= 0)
        c = cl[eo[g]: eo[g + 1]]
        assert c.size == 0 or int(c.max()) < V


def test_batch_graph_is_the_seeded_rmat():
    """Graph g of the batch is R-MAT(scale_g, m_gen_g, seed + g)."""
    hb = pa.HostBatch.rmat(12, 4, 10, seed=100)
    for g in (0, 5, 11):
        one = hb.graph(g)
        ref = pa.HostGraph.rmat(int(hb.scales[g]), hb.m_gen[g] / float(1 << int(hb.scales[g])),
                                seed=100 + g)
        np.testing.assert_array_equal(one.rowptr, ref.rowptr)
        np.testing.assert_array_equal(one.col, ref.col)


def test_batch_scale_draws_cover_range():
    hb = pa.HostBatch.rmat(400, 3, 8, ef_min=1.0, ef_max=2.0, seed=1)
    counts = np.bincount(hb.scales, minlength=9)[3:9]
    assert counts.min() > 40  # uniform over the 6 integers: ~67 each


def test_from_graphs_roundtrip():
    gs = [pa.HostGraph.er(50, 120, seed=s) for s in range(5)]
    gs.append(pa.HostGraph.from_csr(np.zeros(1, np.uint64), np.zeros(0, np.uint32)))
    hb = pa.HostBatch.from_graphs(gs)
    assert hb.G == 6 and hb.vertices == 250
    for g, h in enumerate(gs):
        x = hb.graph(g)
        np.testing.assert_array_equal(x.rowptr, h.rowptr)
        np.testing.assert_array_equal(x.col, h.col)
        np.testing.assert_array_equal(oracle.bz(x.rowptr, x.col), oracle.bz(h.rowptr, h.col))
