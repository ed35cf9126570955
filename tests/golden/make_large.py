"""Write size-independent summaries of the oracle's results on the full-size
workloads (same seeded generators as the product), for GPU tests at sizes
where running the oracle inside a test would take minutes or hours:

* rmat20 / rmat22: triangle counts (parallel MSA restatement) and kmax,
  histogram and a position-weighted checksum of the truss numbers (per
  entry of the symmetric pattern) from the oracle's independent peeling
  (oracle/oracle.cpp, or_truss_peel)                      -> rmat_large.json
* chung_lu: the Chung-Lu graph of the learned-selector evaluation (exponent
  2.0, 4M vertices, mean degree 16, seed 1): the same summaries
                                                           -> rmat_large.json
* grid: the triangulated 2048 x 2048 grid; its closed forms (T = 2(R-1)^2,
  sum of supports 3T, every truss number 3) checked against the oracle
                                                           -> rmat_large.json
* bc20: batched Brandes BC (FP64 restatement) on R-MAT scale 20 for the
  4 x 256 seeded sources of the BC workload: the source list, the sum, and
  exact values at the 256 largest and 4096 seeded random vertices
                                                           -> bc_rmat20.npz
* batch: the 4096-graph batch (the batching workload): every graph's
  triangle count and the number of edges in its 4-truss (the reference's
  k_truss loop), graph by graph                            -> batch4096.npz

    python tests/golden/make_large.py [rmat20 rmat22 chung_lu grid bc20 batch]

No argument writes everything (hours on 8 cores: the sequential peeling
of the scale-22 and Chung-Lu graphs dominates).  A partial run keeps the
other entries of rmat_large.json.
"""
import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

from graphkokkos_amd import HostGraph  # noqa: E402
from graphkokkos_amd.graph import batch_config  # noqa: E402
from oracle import oracle as O  # noqa: E402

RMAT = dict(a=0.57, b=0.19, c=0.19)
GRID_R = 2048


def checksum(a: np.ndarray) -> int:
    w = (np.arange(a.size, dtype=np.uint64) % np.uint64(1000003)) + np.uint64(1)
    return int(np.sum(a.astype(np.uint64) * w, dtype=np.uint64))


def bc_sources(A, nsrc: int = 1024, seed: int = 1) -> np.ndarray:
    """The BC workload's sources: nsrc distinct vertices with at least one
    neighbour, drawn with numpy's PCG64 from `seed`, sorted."""
    rng = np.random.default_rng(seed)
    cand = np.nonzero(np.diff(A.Rst) > 0)[0]
    return np.sort(rng.choice(cand, size=nsrc, replace=False)).astype(np.int32)


def bc_sample_vertices(bc: np.ndarray, seed: int = 7) -> np.ndarray:
    """The 256 vertices with the largest BC plus 4096 seeded random ones."""
    top = np.argsort(-bc, kind="stable")[:256]
    rnd = np.random.default_rng(seed).choice(bc.size, size=4096, replace=False)
    return np.unique(np.concatenate([top, rnd])).astype(np.int64)


def chung_lu_graph():
    return HostGraph.chung_lu(1 << 22, 2.0, 16.0, seed=1)


def batch_graph(cfg: dict, g: int):
    if cfg["kind"][g] == 0:
        return HostGraph.er(int(cfg["n"][g]), float(cfg["mean_degree"][g]), seed=int(cfg["seed"][g]))
    return HostGraph.chung_lu(int(cfg["n"][g]), float(cfg["exponent"][g]), float(cfg["mean_degree"][g]),
                              seed=int(cfg["seed"][g]))


def truss_summary(G) -> dict:
    t0 = time.perf_counter()
    tr, kmax = O.truss_peel(G.A)
    print(f"  peeling {time.perf_counter() - t0:.0f}s, kmax {kmax}", flush=True)
    return {"kmax": kmax, "truss_hist": np.bincount(tr).tolist(), "truss_checksum": checksum(tr)}


def make_rmat(scale: int) -> dict:
    G = HostGraph.rmat(scale, 16, seed=1, **RMAT)
    out = {"edges": G.edges, "triangles": O.tc_sandia(G.L)}
    out.update(truss_summary(G))
    return out


def make_chung_lu() -> dict:
    G = chung_lu_graph()
    out = {"n": G.n, "edges": G.edges, "triangles": O.tc_sandia(G.L)}
    out.update(truss_summary(G))
    return out


def make_grid() -> dict:
    G = HostGraph.grid(GRID_R)
    R = GRID_R
    want = {"edges": 3 * R * R - 4 * R + 1, "triangles": 2 * (R - 1) ** 2}
    assert G.edges == want["edges"]
    assert O.tc_sandia(G.L) == want["triangles"]
    sup = O.support(G.A)
    assert int(sup.sum(dtype=np.int64)) == 6 * want["triangles"]  # both directions of every edge
    tr, kmax = O.truss_peel(G.A)
    assert kmax == 3 and np.all(tr == 3)
    want.update({"support_sum_undirected": 3 * want["triangles"], "kmax": 3})
    return want


def make_bc20() -> None:
    G = HostGraph.rmat(20, 16, seed=1, **RMAT)
    src = bc_sources(G.A)
    t0 = time.perf_counter()
    bc = O.bc_batch(G.A, src)
    print(f"  bc {time.perf_counter() - t0:.0f}s", flush=True)
    vs = bc_sample_vertices(bc)
    np.savez_compressed(HERE / "bc_rmat20.npz", sources=src, vertices=vs, values=bc[vs],
                        total=np.float64(bc.sum()), vmax=np.float64(bc.max()), argmax=np.int64(np.argmax(bc)),
                        edges=np.int64(G.edges))


def make_batch() -> None:
    cfg = batch_config(4096, seed=1)
    ng = len(cfg["n"])
    tri = np.zeros(ng, np.uint64)
    t4 = np.zeros(ng, np.int64)
    edges = np.zeros(ng, np.int64)
    t0 = time.perf_counter()
    for g in range(ng):
        G = batch_graph(cfg, g)
        tri[g] = O.tc_sandia(G.L)
        keep, _ = O.ktruss(G.A, 4)
        t4[g] = int(keep.sum()) // 2
        edges[g] = G.edges
        if g % 256 == 255:
            print(f"  batch {g + 1}/{ng} graphs, {time.perf_counter() - t0:.0f}s", flush=True)
    np.savez_compressed(HERE / "batch4096.npz", triangles=tri, truss4_edges=t4, edges=edges)


def main(argv):
    names = argv or ["rmat20", "rmat22", "chung_lu", "grid", "bc20", "batch"]
    path = HERE / "rmat_large.json"
    out = json.loads(path.read_text()) if path.exists() else {}
    for name in names:
        print(name, flush=True)
        if name in ("rmat20", "rmat22"):
            out[name] = make_rmat(int(name[4:]))
        elif name == "chung_lu":
            out["chung_lu_4M_exp2.0_d16"] = make_chung_lu()
        elif name == "grid":
            out["grid2048"] = make_grid()
        elif name == "bc20":
            make_bc20()
        elif name == "batch":
            make_batch()
        else:
            raise SystemExit(f"unknown fixture {name}")
        path.write_text(json.dumps(out))


if __name__ == "__main__":
    main(sys.argv[1:])
