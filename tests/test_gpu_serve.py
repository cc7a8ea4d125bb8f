"""GPU tests of the request loop (kc_server) through the C ABI: a burst of
fresh-graph decompositions of mixed sizes (small ones fused into batched
launches) and in-order edge updates on a resident graph, overlapping on
several streams.  Every result must equal the same request served alone,
and the oracle's Batagelj-Zaversnik cores (via the core checksum).
"""
import numpy as np
import pytest

import oracle
import pico_amd as pa

pytestmark = pytest.mark.gpu


def _mix(rng, n_dec, n_upd, V):
    reqs = []
    for i in range(n_dec):
        reqs.append(("rmat", int(rng.integers(6, 17)), float(rng.uniform(2, 16)), 1000 + i))
    for i in range(n_upd):
        op = "insert" if i % 2 == 0 else "delete"
        u = rng.integers(0, V, 400, dtype=np.uint32)
        v = rng.integers(0, V, 400, dtype=np.uint32)
        reqs.append(("update", op, u, v))
    order = rng.permutation(len(reqs))
    return [reqs[i] for i in order]


def _serve(srv, gid, reqs, one_at_a_time):
    out = {}
    pending = 0
    for k, r in enumerate(reqs):
        if r[0] == "rmat":
            srv.submit_rmat(r[1], r[2], seed=r[3], tag=k)
        else:
            srv.submit_update(gid, r[1], r[2], r[3], tag=k)
        pending += 1
        if one_at_a_time:
            for c in srv.poll(wait_ms=-1):
                out[c["tag"]] = c
                pending -= 1
    while pending:
        for c in srv.poll(wait_ms=-1):
            out[c["tag"]] = c
            pending -= 1
    return out


def test_serve_mix_equals_alone_and_oracle():
    rng = np.random.default_rng(7)
    scale_base = 12
    reqs = _mix(rng, 40, 12, 1 << scale_base)
    srv = pa.Server(slots=4, max_batch=16, fuse_max_scale=14)
    gid = srv.gen_rmat_graph(scale_base, 8.0, seed=5)
    mixed = _serve(srv, gid, reqs, one_at_a_time=False)
    st = srv.stats()
    final_mixed = srv.download_graph(gid)
    srv.close()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["fused_jobs"] >= 1  # small decompositions shared launches

    alone = pa.Server(slots=1, max_batch=1)
    gid2 = alone.gen_rmat_graph(scale_base, 8.0, seed=5)
    ref = _serve(alone, gid2, reqs, one_at_a_time=True)
    alone.close()
    for k in range(len(reqs)):
        assert mixed[k]["status"] == 0 and ref[k]["status"] == 0, k
        assert mixed[k]["checksum"] == ref[k]["checksum"], (k, reqs[k][:2])
    # fresh decompositions against the oracle
    for k, r in enumerate(reqs):
        if r[0] != "rmat" or k % 3:
            continue
        hg = pa.HostGraph.rmat(r[1], r[2], seed=r[3])
        want = oracle.bz(hg.rowptr, hg.col)
        assert mixed[k]["checksum"] == pa.core_checksum(want), k
        assert mixed[k]["kmax"] == int(want.max())
    # the resident graph after all updates: the last update's cores
    last = max(k for k, r in enumerate(reqs) if r[0] == "update")
    want = oracle.bz(final_mixed.rowptr, final_mixed.col)
    assert mixed[last]["checksum"] == pa.core_checksum(want)


def test_serve_core_out_and_admission():
    """A budget that admits one large request at a time: requests queue and
    still all complete, with cores delivered into the callers' buffers."""
    srv = pa.Server(slots=3, hbm_budget=30 << 20, fuse_max_scale=8)
    bufs = []
    for i, scale in enumerate([13, 14, 12, 13]):
        b = np.zeros(1 << scale, np.uint32)
        bufs.append((scale, i, b))
        srv.submit_rmat(scale, 8.0, seed=50 + i, tag=i, core_out=b)
    got = {}
    while len(got) < 4:
        for c in srv.poll(wait_ms=-1):
            got[c["tag"]] = c
    st = srv.stats()
    srv.close()
    assert st["max_inflight_bytes"] <= 30 << 20
    for scale, i, b in bufs:
        hg = pa.HostGraph.rmat(scale, 8.0, seed=50 + i)
        np.testing.assert_array_equal(b, oracle.bz(hg.rowptr, hg.col))
        assert got[i]["checksum"] == pa.core_checksum(b)


def test_serve_rejects_oversized_and_bad_requests():
    srv = pa.Server(slots=1, hbm_budget=1 << 20)
    with pytest.raises(pa.KCoreError, match="budget"):
        srv.submit_rmat(16, 16.0, seed=1)
    with pytest.raises(pa.KCoreError):
        srv.submit_update(3, "insert", [0], [1])  # no such graph
    srv.close()
