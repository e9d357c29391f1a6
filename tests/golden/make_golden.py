"""Writes the golden vectors tests/golden/<graph>.npz.

For each of the reference's METIS fixtures (graphs/test1-6, jazz; copied to
tests/golden/graphs) and for ER G(198, 2742) seed 0, the CPU oracle's
distances, path counts, float dependencies and FP64 BC over all sources are
stored together with the graph (R, C).  Before anything is written the
values are cross-checked against scipy's BFS and networkx's Brandes, so the
fixtures are not only the oracle's own output.

Run from the repository root after build():  python tests/golden/make_golden.py
"""
import os
import sys
import tempfile

import networkx as nx
import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import shortest_path

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import __graft_entry__  # noqa: E402
from oracle import oracle as orc  # noqa: E402

INT_MAX = 2147483647


def graphs():
    gu = __graft_entry__.load_package()
    for name in ["test1", "test2", "test3", "test4", "test5", "test6", "jazz"]:
        p = os.path.join(HERE, "graphs", name + ".graph")
        if orc.ref_parse_available():
            R, C, _, _, _, _ = orc.ref_parse_metis(p)
        else:
            hg = gu.HostGraph.parse_metis(p)
            R, C = hg.R, hg.C
        yield name, R, C
    hg = gu.HostGraph.er(198, 2742, seed=0)
    with tempfile.TemporaryDirectory() as t:
        p = os.path.join(t, "er.graph")
        hg.write_metis(p)
        hg = gu.HostGraph.parse_metis(p)
    yield "er198", hg.R, hg.C


def main():
    for name, R, C in graphs():
        n = len(R) - 1
        dist = np.empty((n, n), np.int32)
        sig = np.empty((n, n), np.uint64)
        delta = np.empty((n, n), np.float32)
        for s in range(n):
            dist[s], sig[s] = orc.sigma(R, C, s)
            delta[s] = orc.brandes_f32(R, C, s)
        bc = orc.bc_f64(R, C, np.arange(n))
        # independent cross-checks
        A = sp.csr_matrix((np.ones(len(C)), C, R), shape=(n, n))
        D = shortest_path(A, unweighted=True, directed=True)
        assert np.array_equal(np.where(np.isinf(D), INT_MAX, D).astype(np.int64), dist.astype(np.int64))
        g = nx.Graph()
        g.add_nodes_from(range(n))
        g.add_edges_from((v, int(C[k])) for v in range(n) for k in range(R[v], R[v + 1]))
        ref = nx.betweenness_centrality(g, normalized=False)
        assert np.allclose(bc, [2 * ref[v] for v in range(n)], rtol=1e-12, atol=1e-9)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), R=R, C=C, dist=dist, sigma=sig,
                            delta_f32=delta, bc_f64=bc)
        print("wrote %s.npz (n=%d, m=%d)" % (name, n, len(C)))


if __name__ == "__main__":
    main()
