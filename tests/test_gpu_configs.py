"""The benchmark configurations at their full sizes, every method against
the oracle's Batagelj-Zaversnik cores (bit-exact):

* R-MAT scale 22, edge factor 16 (4.19M vertices, 64.2M edges), int32 CSR;
* heavy-skew Chung-Lu, n = 2^24, exponent 2.1, mean degree 20, expected
  degrees capped at 4M, int64 rowptr: hub vertices of up to ~10^5 neighbours
  go through the virtual-node split of both engine families.

Both graphs are built on the device (the host generators' graphs, edge for
edge) and copied back for the oracle.
"""
import time

import numpy as np
import pytest

import oracle
import pico_amd as pa

pytestmark = pytest.mark.gpu

METHODS = ["peelone_dyn", "peel_dyn", "gpp", "peelone", "histocore", "cntcore", "nbrcore"]


@pytest.fixture(scope="module")
def ctx():
    c = pa.Context(0)
    yield c
    c.close()


def _all_methods(dg, hg, label):
    t = time.time()
    want = oracle.bz(hg.rowptr, hg.col)
    print("[%s] V=%d nnz=%d kmax=%d oracle %.1f s" % (label, hg.V, hg.nnz, want.max(),
                                                      time.time() - t), flush=True)
    for m in METHODS:
        core, st = dg.decompose(m)
        print("[%s] %-12s %.2f ms" % (label, m, st["device_ms"]), flush=True)
        np.testing.assert_array_equal(core, want, err_msg="%s %s" % (label, m))
        assert st["kmax"] == int(want.max())


def test_rmat22_every_method(ctx):
    dg = ctx.gen_rmat(22, 16, seed=1)
    hg = dg.download()
    assert hg.V == 1 << 22 and hg.nnz == 2 * 64151999
    _all_methods(dg, hg, "rmat22")
    dg.close()


def test_chunglu_2e24_every_method(ctx):
    dg = ctx.gen_chunglu(1 << 24, beta=2.1, mean_deg=20.0, cap=4e6, seed=3, rowptr_bytes=8)
    hg = dg.download()
    deg = np.diff(hg.rowptr)
    print("[chunglu] max degree %d" % deg.max(), flush=True)
    assert hg.V == 1 << 24 and deg.max() > 4096  # hubs split into virtual nodes
    _all_methods(dg, hg, "chunglu")
    dg.close()
