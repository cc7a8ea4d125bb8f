"""The oracle (CPU restatement) pinned against what fixes its answers
independently: networkx.core_number fixtures, closed forms, brute force on
tiny graphs, hand-derived iteration counts on example.txt, and cross-method
agreement.  No GPU."""
import os

import numpy as np
import pytest

import oracle
from tests import graphs

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cores.npz")


def _golden():
    z = np.load(GOLD)
    names = sorted({k.split("__")[0] for k in z.files})
    return [(n, z[n + "__rowptr"], z[n + "__col"], z[n + "__core"]) for n in names]


GOLDEN = _golden()


@pytest.mark.parametrize("name,rp,col,core", GOLDEN, ids=[g[0] for g in GOLDEN])
def test_all_methods_match_networkx(name, rp, col, core):
    np.testing.assert_array_equal(oracle.bz(rp, col), core)
    np.testing.assert_array_equal(oracle.gpp(rp, col, compat=False)[0], core)
    np.testing.assert_array_equal(oracle.peelone(rp, col)[0], core)
    np.testing.assert_array_equal(oracle.bucket_peel(rp, col)[0], core)
    for fn in (oracle.histocore, oracle.cntcore, oracle.nbrcore, oracle.hindex_jacobi):
        np.testing.assert_array_equal(fn(rp, col)[0], core)
    assert oracle.check_invariants(rp, col, core) == 0


@pytest.mark.parametrize("name,rp,col,core", GOLDEN, ids=[g[0] for g in GOLDEN])
def test_hindex_round_counts_equal_jacobi(name, rp, col, core):
    # the three frontier schemes are Jacobi iterations of c <- min(c, H(c));
    # their round counts must equal the plain synchronous iteration's
    _, jac = oracle.hindex_jacobi(rp, col)
    assert oracle.histocore(rp, col)[1] == jac
    assert oracle.cntcore(rp, col)[1] == jac
    assert oracle.nbrcore(rp, col)[1] == jac


@pytest.mark.parametrize("name,rp,col,core", GOLDEN, ids=[g[0] for g in GOLDEN])
def test_bucket_levels_and_gpp_quirk(name, rp, col, core):
    _, levels = oracle.bucket_peel(rp, col)
    assert levels == int(core.max()) + 1 if core.size else levels == 0
    compat, _ = oracle.gpp(rp, col, compat=True)
    iso = np.diff(rp) == 0
    np.testing.assert_array_equal(compat[iso], 1)       # gppeel.cu:182-191
    np.testing.assert_array_equal(compat[~iso], core[~iso])


@pytest.mark.parametrize("name", graphs.CLOSED_FORMS)
def test_closed_forms_and_bruteforce(name):
    rp, col, expect = graphs.closed_form(name)
    np.testing.assert_array_equal(oracle.bz(rp, col), expect)
    np.testing.assert_array_equal(oracle.bruteforce(rp, col), expect)


def test_bruteforce_random_tiny():
    rng = np.random.default_rng(0)
    for _ in range(60):
        n = int(rng.integers(2, 30))
        m = int(rng.integers(0, n * (n - 1) // 2 + 1))
        edges = set()
        while len(edges) < m:
            u, v = rng.integers(0, n, 2)
            if u != v:
                edges.add((min(u, v), max(u, v)))
        rp, col = graphs.csr(n, sorted(edges))
        bf = oracle.bruteforce(rp, col)
        np.testing.assert_array_equal(oracle.bz(rp, col), bf)
        np.testing.assert_array_equal(oracle.histocore(rp, col)[0], bf)
        np.testing.assert_array_equal(oracle.gpp(rp, col, compat=False)[0], bf)


def test_example_txt_hand_derived_counts():
    # Derived by hand from the reference loops on example.txt (5 vertices,
    # every core 2):
    #  gppeel  : L1 empty scan; L2: {4} / {1,5} / {2,3} retire; final empty
    #            scan -> 5 iterations
    #  h-index : round 1 lowers {1,5} to 2, round 2 lowers {2,3}, round 3 sees
    #            no change -> 3 iterations (HistoCore, CntCore, NbrCore)
    #  bucket  : levels 0..2 -> 3 loop trips, kmax 2
    rp, col = graphs.example()
    assert oracle.gpp(rp, col)[1] == 5
    for fn in (oracle.histocore, oracle.cntcore, oracle.nbrcore, oracle.hindex_jacobi):
        c, it = fn(rp, col)
        assert c.tolist() == [2] * 5 and it == 3
    c, lv = oracle.bucket_peel(rp, col)
    assert c.tolist() == [2] * 5 and lv == 3


def test_example_txt_histocore_final_histogram():
    # final bins, hand-derived: vertex slots hold the old core of vertices
    # that changed (3) or the degree; collapsed top bins hold the suffix sums
    rp, col = graphs.example()
    _, _, hist = oracle.histocore(rp, col, want_hist=True)
    assert hist.tolist() == [3, 0, 3, 2,  3, 0, 3, 1,  3, 0, 3, 1,  2, 0, 2,  3, 0, 3, 2]


def test_tiny_gpp_iteration_counts():
    # path 0-1-2: L1 retires {0,2}, then {1}, then an empty final scan -> 3
    rp, col = graphs.csr(3, [(0, 1), (1, 2)])
    assert oracle.gpp(rp, col)[1] == 3
    # triangle: L1 empty scan, L2 retires all, final empty scan -> 3
    rp, col = graphs.csr(3, [(0, 1), (1, 2), (0, 2)])
    assert oracle.gpp(rp, col)[1] == 3
    # one isolated vertex: retired at level 1 with core 1, then final -> 2
    rp, col = graphs.csr(1, [])
    c, it = oracle.gpp(rp, col, compat=True)
    assert it == 2 and c.tolist() == [1]


def test_invariant_check_catches_plausible_mistakes():
    hg_rp, hg_col = graphs.example()
    good = oracle.bz(hg_rp, hg_col)
    assert oracle.check_invariants(hg_rp, hg_col, good) == 0
    bad = good.copy()
    bad[0] += 1                      # off by one upwards
    assert oracle.check_invariants(hg_rp, hg_col, bad) != 0
    bad = good.copy()
    bad[3] = 1                       # one too low breaks the fixed point
    assert oracle.check_invariants(hg_rp, hg_col, bad) != 0
