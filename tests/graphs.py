"""Small test graphs with known core numbers (closed forms from the
definition of the k-core) and the reference's example.txt fixture."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def csr(n, edges):
    rows = [[] for _ in range(n)]
    for u, v in edges:
        if u == v:
            continue
        rows[u].append(v)
        rows[v].append(u)
    rp = np.zeros(n + 1, np.uint64)
    rp[1:] = np.cumsum([len(r) for r in rows])
    col = np.array([x for r in rows for x in sorted(r)], np.uint32)
    return rp, col


def example():
    """example.txt (K5 minus {2,4},{3,4},{1,5}): every vertex has core 2."""
    edges = []
    with open(os.path.join(GOLDEN, "example.txt")) as f:
        lines = f.read().splitlines()[2:]
    for ln in lines:
        a, b = ln.split()[:2]
        edges.append((int(a) - 1, int(b) - 1))
    return csr(5, edges)


def _complete(n, off=0):
    return [(off + i, off + j) for i in range(n) for j in range(i + 1, n)]


def closed_form(name):
    """(rowptr, col, expected core) for graphs whose cores follow directly
    from the k-core definition."""
    if name == "k1":
        rp, col = csr(1, [])
        return rp, col, np.array([0], np.uint32)
    if name == "k2":
        rp, col = csr(2, [(0, 1)])
        return rp, col, np.array([1, 1], np.uint32)
    if name == "k7":  # K_n -> n-1
        rp, col = csr(7, _complete(7))
        return rp, col, np.full(7, 6, np.uint32)
    if name == "cycle9":  # cycle -> 2
        rp, col = csr(9, [(i, (i + 1) % 9) for i in range(9)])
        return rp, col, np.full(9, 2, np.uint32)
    if name == "star12":  # star -> 1
        rp, col = csr(12, [(0, i) for i in range(1, 12)])
        return rp, col, np.full(12, 1, np.uint32)
    if name == "petersen":  # 3-regular connected -> 3
        outer = [(i, (i + 1) % 5) for i in range(5)]
        spokes = [(i, i + 5) for i in range(5)]
        inner = [(5 + i, 5 + (i + 2) % 5) for i in range(5)]
        rp, col = csr(10, outer + spokes + inner)
        return rp, col, np.full(10, 3, np.uint32)
    if name == "empty5":  # isolated -> 0
        rp, col = csr(5, [])
        return rp, col, np.zeros(5, np.uint32)
    if name == "forest":  # forest -> 1 on non-isolated, 0 isolated
        edges = [(0, 1), (1, 2), (1, 3), (3, 4), (5, 6), (7, 8), (8, 9), (8, 10)]
        rp, col = csr(12, edges)
        return rp, col, np.array([1] * 11 + [0], np.uint32)
    if name == "k33":  # K_{a,b} -> min(a,b)
        rp, col = csr(6, [(i, 3 + j) for i in range(3) for j in range(3)])
        return rp, col, np.full(6, 3, np.uint32)
    if name == "two_cliques":  # K4 + K6 joined by a bridge: 3 and 5
        edges = _complete(4) + _complete(6, 4) + [(0, 4)]
        rp, col = csr(10, edges)
        return rp, col, np.array([3] * 4 + [5] * 6, np.uint32)
    raise KeyError(name)


CLOSED_FORMS = ["k1", "k2", "k7", "cycle9", "star12", "petersen", "empty5", "forest", "k33",
                "two_cliques"]
