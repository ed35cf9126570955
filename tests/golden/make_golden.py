"""Write the golden fixtures under tests/golden/ with library routines only
(networkx 3.4 and scipy), never with the oracle or the product:

* triangles:  sum(nx.triangles(G).values()) / 3
* supports:   ((A @ A) .* A) with scipy sparse integers, per stored entry
* truss:      largest k with the edge in nx.k_truss(G, k)
* bc:         2 * nx.betweenness_centrality_subset(G, sources, all targets,
              normalized=False)  -- networkx halves undirected subset scores,
              the masked-product formulation counts every source's delta once

Graphs: Zachary's karate club, a Petersen graph, a 5x5 triangulated grid, an
Erdos-Renyi graph (nx.gnm_random_graph, seed 7) and a small power-law graph
(nx.powerlaw_cluster_graph, seed 3).  Each fixture is an .npz of plain
arrays (no pickles): n, the edge list, and the expected outputs on the
symmetric CSR with sorted rows built from that edge list.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

from pathlib import Path

import networkx as nx
import numpy as np
import scipy.sparse as sp

HERE = Path(__file__).resolve().parent


def grid_diag(r: int) -> nx.Graph:
    G = nx.Graph()
    G.add_nodes_from(range(r * r))
    for i in range(r):
        for j in range(r):
            v = i * r + j
            if j + 1 < r:
                G.add_edge(v, v + 1)
            if i + 1 < r:
                G.add_edge(v, v + r)
            if i + 1 < r and j + 1 < r:
                G.add_edge(v, v + r + 1)
    return G


def graphs() -> dict[str, nx.Graph]:
    return {
        "karate": nx.karate_club_graph(),
        "petersen": nx.petersen_graph(),
        "grid5": grid_diag(5),
        "er60": nx.gnm_random_graph(60, 240, seed=7),
        "plc80": nx.powerlaw_cluster_graph(80, 4, 0.6, seed=3),
    }


def csr_of(G: nx.Graph):
    n = G.number_of_nodes()
    nodes = sorted(G.nodes())
    assert nodes == list(range(n))
    A = nx.to_scipy_sparse_array(G, nodelist=nodes, dtype=np.int64, format="csr")
    A.data[:] = 1
    A.sort_indices()
    return n, A


def truss_numbers(G: nx.Graph, A) -> np.ndarray:
    t = {tuple(sorted(e)): 2 for e in G.edges()}
    k = 3
    while True:
        H = nx.k_truss(G, k)
        if H.number_of_edges() == 0:
            break
        for e in H.edges():
            t[tuple(sorted(e))] = k
        k += 1
    out = np.zeros(A.nnz, np.uint32)
    for i in range(A.shape[0]):
        for p in range(A.indptr[i], A.indptr[i + 1]):
            j = int(A.indices[p])
            out[p] = t[(min(i, j), max(i, j))]
    return out


def main() -> None:
    rng = np.random.default_rng(11)
    for name, G in graphs().items():
        G = nx.convert_node_labels_to_integers(G)
        n, A = csr_of(G)
        S = (A @ A).multiply(A).tocsr()
        S.sort_indices()
        # S has A's pattern where the support is nonzero; expand to A's entries
        sup = np.zeros(A.nnz, np.int32)
        for i in range(n):
            row = dict(zip(S.indices[S.indptr[i]:S.indptr[i + 1]], S.data[S.indptr[i]:S.indptr[i + 1]]))
            for p in range(A.indptr[i], A.indptr[i + 1]):
                sup[p] = row.get(int(A.indices[p]), 0)
        tri = sum(nx.triangles(G).values()) // 3
        sources = np.sort(rng.choice(n, size=min(n, 12), replace=False)).astype(np.int32)
        bcs = nx.betweenness_centrality_subset(G, sources=list(map(int, sources)), targets=list(G.nodes()),
                                               normalized=False)
        bc = np.array([2.0 * bcs[v] for v in range(n)])
        edges = np.array(sorted(tuple(sorted(e)) for e in G.edges()), np.int32).reshape(-1, 2)
        np.savez(HERE / f"{name}.npz", n=np.int64(n), edges=edges, rst=A.indptr.astype(np.int32),
                 col=A.indices.astype(np.int32), triangles=np.int64(tri), support=sup,
                 truss=truss_numbers(G, A), sources=sources, bc=bc)
        print(f"{name}: n={n} m={G.number_of_edges()} T={tri}")


if __name__ == "__main__":
    main()
