"""Writes the committed golden fixtures (run from the repo root, here, where
networkx and the reference tree are available; the fixtures are data):

  tests/golden/cores.npz          networkx.core_number on small graphs
  tests/golden/reader_golden.npz  layouts produced by the reference's own host
                                  readers (oracle/_ref, built from the
                                  reference sources) on tests/golden/readers/*

    python tests/golden/make_golden.py
"""
import os
import sys
import tempfile

import networkx as nx
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def csr_of(G, n):
    rows = [[] for _ in range(n)]
    for u, v in G.edges():
        if u == v:
            continue
        rows[u].append(v)
        rows[v].append(u)
    rp = np.zeros(n + 1, np.uint64)
    rp[1:] = np.cumsum([len(r) for r in rows])
    col = np.array([x for r in rows for x in sorted(r)], np.uint32)
    return rp, col


def graphs():
    yield "er_200_600", nx.gnm_random_graph(200, 600, seed=1)
    yield "er_2000_9000", nx.gnm_random_graph(2000, 9000, seed=2)
    yield "ba_1500_5", nx.barabasi_albert_graph(1500, 5, seed=3)
    yield "ws_1000_8", nx.watts_strogatz_graph(1000, 8, 0.2, seed=4)
    yield "powerlaw_cluster_1200", nx.powerlaw_cluster_graph(1200, 4, 0.3, seed=5)
    yield "karate", nx.karate_club_graph()
    yield "les_mis", nx.convert_node_labels_to_integers(nx.les_miserables_graph())
    yield "florentine", nx.convert_node_labels_to_integers(nx.florentine_families_graph())
    yield "grid_20x30", nx.convert_node_labels_to_integers(nx.grid_2d_graph(20, 30))
    yield "lollipop_12_20", nx.lollipop_graph(12, 20)
    yield "caveman_6_7", nx.connected_caveman_graph(6, 7)
    yield "tree_3_6", nx.balanced_tree(3, 6)
    G = nx.gnm_random_graph(600, 1500, seed=6)
    G.add_nodes_from(range(600, 640))  # isolated tail
    yield "er_with_isolated", G
    yield "chung_lu_1500", nx.expected_degree_graph(
        [min(300.0, 400.0 * (i + 1) ** -0.9) + 1.0 for i in range(1500)], seed=7, selfloops=False)


def make_cores():
    out = {}
    for name, G in graphs():
        n = G.number_of_nodes()
        G = nx.Graph(G)
        G.remove_edges_from(nx.selfloop_edges(G))
        rp, col = csr_of(G, n)
        cn = nx.core_number(G)
        out[name + "__rowptr"] = rp
        out[name + "__col"] = col
        out[name + "__core"] = np.array([cn[i] for i in range(n)], np.uint32)
    np.savez_compressed(os.path.join(HERE, "cores.npz"), **out)
    print("cores.npz:", len(out) // 3, "graphs")


def make_readers():
    import oracle
    if not oracle.ref_available():
        oracle.build(ref=True)
    rdir = os.path.join(HERE, "readers")
    out = {}
    files = sorted(f for f in os.listdir(rdir) if f.endswith(".txt"))
    files.append("../example.txt")
    for f in files:
        path = os.path.join(rdir, f)
        key = os.path.basename(f)[:-4]
        with tempfile.TemporaryDirectory() as wd:
            coo = oracle.ref_read_coo(path, wd)
            out[key + "__coo_meta"] = np.array([coo["num_nodes"], coo["num_edges"],
                                                int(coo["has_zero"])], np.uint64)
            out[key + "__coo_node_pointer"] = coo["node_pointer"]
            out[key + "__coo_edge_list"] = coo["edge_list"]
            out[key + "__coo_parts"] = coo["parts"].reshape(-1)
        with tempfile.TemporaryDirectory() as wd:
            try:
                csr = oracle.ref_read_csr(path, wd)
                out[key + "__csr_meta"] = np.array([csr["V"], csr["E"]], np.uint64)
                out[key + "__csr_degrees"] = csr["degrees"]
                out[key + "__csr_offsets"] = csr["offsets"]
                out[key + "__csr_neighbors"] = csr["neighbors"]
            except Exception as ex:  # 0-based input wraps ids: the reference fails
                out[key + "__csr_failed"] = np.array([1], np.uint8)
                print("csr reader failed on", f, type(ex).__name__)
    np.savez_compressed(os.path.join(HERE, "reader_golden.npz"), **out)
    print("reader_golden.npz:", len(files), "files")


if __name__ == "__main__":
    make_cores()
    make_readers()
