"""GPU parity of core maintenance under batched edge insertions and deletions
(kc_update) through the C ABI.  The reference has no dynamic path; the
oracle is a full Batagelj-Zaversnik recomputation on the updated graph,
kept on the host beside the device graph.  Core numbers are compared bit
for bit after every batch, and kc_decompose on the updated device graph
must agree as well.
"""
import numpy as np
import pytest

import oracle
import pico_amd as pa

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = pa.Context(0)
    yield c
    c.close()


class EdgeSet:
    """Host mirror of the graph as a set of unordered pairs."""

    def __init__(self, hg):
        rp, col = hg.rowptr, hg.col
        src = np.repeat(np.arange(hg.V, dtype=np.uint64), np.diff(rp).astype(np.int64))
        keep = src < col
        self.V = hg.V
        self.keys = set((src[keep] << np.uint64(32) | col[keep].astype(np.uint64)).tolist())

    def apply(self, op, u, v):
        for a, b in zip(u.tolist(), v.tolist()):
            if a == b:
                continue
            k = (min(a, b) << 32) | max(a, b)
            if op == "insert":
                self.keys.add(k)
            else:
                self.keys.discard(k)

    def csr(self):
        k = np.array(sorted(self.keys), np.uint64)
        a = (k >> np.uint64(32)).astype(np.uint32)
        b = (k & np.uint64(0xffffffff)).astype(np.uint32)
        hg = pa.HostGraph.from_pairs(self.V, a, b)
        return hg.rowptr, hg.col

    def sample_present(self, rng, n):
        arr = np.array(list(self.keys), np.uint64)
        pick = arr[rng.choice(arr.size, size=min(n, arr.size), replace=False)]
        return ((pick >> np.uint64(32)).astype(np.uint32),
                (pick & np.uint64(0xffffffff)).astype(np.uint32))


def _run(ctx, hg, batches, seed, width=4):
    rng = np.random.default_rng(seed)
    es = EdgeSet(hg)
    dg = ctx.upload(hg, rowptr_bytes=width)
    V = hg.V
    for i, (op, n) in enumerate(batches):
        if op == "insert":
            u = rng.integers(0, V, n, dtype=np.uint32)
            v = rng.integers(0, V, n, dtype=np.uint32)
        else:
            u, v = es.sample_present(rng, n)
            if n > 4:  # plus a few absent pairs, which are skipped
                u = np.concatenate([u, rng.integers(0, V, 4, dtype=np.uint32)])
                v = np.concatenate([v, rng.integers(0, V, 4, dtype=np.uint32)])
        before = len(es.keys)
        es.apply(op, u, v)
        core, st = dg.update(op, u, v)
        rp, col = es.csr()
        want = oracle.bz(rp, col)
        np.testing.assert_array_equal(core, want, err_msg="batch %d (%s %d)" % (i, op, n))
        assert st["applied"] == abs(len(es.keys) - before), (i, st)
        assert dg.nnz == 2 * len(es.keys)
    # the updated device graph decomposes to the same cores
    for method in ("peelone_dyn", "histocore", "gpp"):
        c2, _ = dg.decompose(method)
        np.testing.assert_array_equal(c2, want, err_msg=method)
    d = dg.download()
    np.testing.assert_array_equal(d.rowptr, rp)
    np.testing.assert_array_equal(np.sort(d.col), np.sort(col))
    dg.close()


def test_update_rmat_mixed(ctx):
    hg = pa.HostGraph.rmat(12, 8, seed=3)
    batches = [("insert", 300), ("delete", 300), ("insert", 1000), ("delete", 1500),
               ("insert", 50), ("delete", 10), ("insert", 3000), ("delete", 3000)]
    _run(ctx, hg, batches, seed=1)


def test_update_er_wide_rowptr(ctx):
    hg = pa.HostGraph.er(3000, 9000, seed=2)
    _run(ctx, hg, [("insert", 2000), ("delete", 2500), ("insert", 400)], seed=2, width=8)


def test_update_hub_growth(ctx):
    """Many insertions at one vertex: one matching each, the row outgrows its
    spare slots and is laid out again."""
    hg = pa.HostGraph.er(2000, 4000, seed=5)
    es = EdgeSet(hg)
    dg = ctx.upload(hg)
    u = np.zeros(600, np.uint32)
    v = np.arange(1, 601, dtype=np.uint32)
    es.apply("insert", u, v)
    core, st = dg.update("insert", u, v)
    assert st["grows"] >= 1 and st["matchings"] == 600  # all share vertex 0
    rp, col = es.csr()
    np.testing.assert_array_equal(core, oracle.bz(rp, col))
    # a clique on 40 vertices: 39 matchings, cores rise to 39
    a, b = np.triu_indices(40, 1)
    a = (a + 1000).astype(np.uint32)
    b = (b + 1000).astype(np.uint32)
    es.apply("insert", a, b)
    core, st = dg.update("insert", a, b)
    rp, col = es.csr()
    want = oracle.bz(rp, col)
    np.testing.assert_array_equal(core, want)
    assert want[1000:1040].min() >= 39 and st["matchings"] >= 39
    # and take it apart again
    es.apply("delete", a, b)
    core, st = dg.update("delete", a, b)
    rp, col = es.csr()
    np.testing.assert_array_equal(core, oracle.bz(rp, col))
    # deletions go in one launch, however many share a vertex: the hub's 600
    # new edges plus some it had before, in scrambled order
    old = col[rp[0]:rp[1]]
    dv = np.random.default_rng(8).permutation(old).astype(np.uint32)
    du = np.zeros_like(dv)
    es.apply("delete", dv, du)
    core, st = dg.update("delete", dv, du)
    assert st["matchings"] == 1 and st["applied"] == len(dv)
    rp, col = es.csr()
    assert rp[1] == rp[0]
    np.testing.assert_array_equal(core, oracle.bz(rp, col))
    dg.close()


def test_update_skips_and_errors(ctx):
    hg = pa.HostGraph.from_csr(*_path(6))
    dg = ctx.upload(hg)
    # self-loop, duplicate pair (both orientations), present edge
    core, st = dg.update("insert", [2, 0, 5, 0], [2, 5, 0, 1])
    assert st["requested"] == 4 and st["applied"] == 1 and st["skipped"] == 3
    assert core.tolist() == [2, 2, 2, 2, 2, 2]  # the path closed into a cycle
    core, st = dg.update("delete", [0, 3], [3, 0])  # absent edge
    assert st["applied"] == 0 and core.tolist() == [2] * 6
    core, st = dg.update("delete", [], [])
    assert st["applied"] == 0
    with pytest.raises(pa.KCoreError, match="out of range"):
        dg.update("insert", [0], [6])
    core, st = dg.update("delete", [4], [5])
    assert core.tolist() == [1] * 6
    dg.close()


def _path(n):
    rp = [0]
    col = []
    for i in range(n):
        nb = [j for j in (i - 1, i + 1) if 0 <= j < n]
        col += nb
        rp.append(len(col))
    return np.array(rp, np.uint64), np.array(col, np.uint32)


def test_update_isolated_and_empty(ctx):
    hg = pa.HostGraph.from_csr(np.zeros(11, np.uint64), np.zeros(0, np.uint32))
    es = EdgeSet(hg)
    dg = ctx.upload(hg)
    rng = np.random.default_rng(9)
    for op in ("insert", "insert", "delete", "insert", "delete"):
        if op == "insert":
            u = rng.integers(0, 10, 12, dtype=np.uint32)
            v = rng.integers(0, 10, 12, dtype=np.uint32)
        else:
            u, v = es.sample_present(rng, 5)
        es.apply(op, u, v)
        core, _ = dg.update(op, u, v)
        rp, col = es.csr()
        np.testing.assert_array_equal(core, oracle.bz(rp, col))
    dg.close()


def test_update_prepare_then_update(ctx):
    hg = pa.HostGraph.rmat(11, 8, seed=12)
    es = EdgeSet(hg)
    dg = ctx.upload(hg)
    dg.update_prepare()
    dg.update_prepare()  # idempotent
    rng = np.random.default_rng(13)
    u = rng.integers(0, hg.V, 300).astype(np.uint32)
    v = rng.integers(0, hg.V, 300).astype(np.uint32)
    es.apply("insert", u, v)
    core, st = dg.update("insert", u, v)
    rp, col = es.csr()
    np.testing.assert_array_equal(core, oracle.bz(rp, col))
    dg.close()
