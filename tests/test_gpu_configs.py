"""GPU parity on the full-size workloads, against summaries of the oracle's
results written by tests/golden/make_large.py (the oracle takes minutes to
hours on these, so it does not run inside the tests):

* BC: 4 x 256 seeded sources on R-MAT scale 20, FP64 (GraphKokkos
  include/betweenness_centrality.h:406-506 over an explicit source list):
  the sum of all scores and exact values at 4352 vertices (the 256 largest
  and 4096 seeded random ones) within 1e-9 relative;
* batching: the 4096 seeded uneven graphs in one captured step, every
  graph's triangle count (src/main.cpp:69-90) and 4-truss edge count
  (src/main.cpp:93-141); the union has 1.1e9 CSR entries, beyond the 2^30
  where int32 bisection midpoints over CSR positions overflow;
* the learned selector's graphs: R-MAT scale 22, the triangulated
  2048 x 2048 grid (closed forms: T = 2(R-1)^2 = 8,380,418, supports summing
  to 3T = 25,141,254 per direction, every truss number 3) and the Chung-Lu
  graph (exponent 2.0, 4M vertices, mean degree 16): triangle counts,
  supports and full truss decompositions with a trained network choosing
  the accumulators (GK_SEL_LEARNED), and with the fixed-threshold fallback.
"""
import json

import numpy as np
import pytest

from conftest import GOLDEN
from graphkokkos_amd import SEL_LEARNED, HostGraph, selector
from graphkokkos_amd._lib import check, lib
from graphkokkos_amd.graph import batch_config

pytestmark = pytest.mark.gpu

RMAT = dict(a=0.57, b=0.19, c=0.19)
BC_RTOL = 1e-9
LEARNED = {"selector": SEL_LEARNED}


def say(*a):
    print(*a, flush=True)


def large():
    return json.loads((GOLDEN / "rmat_large.json").read_text())


def summary_of(key):
    """The oracle's summary of one graph; its truss histogram is written
    last (the sequential peeling takes up to hours), and the test waits for
    it rather than checking less."""
    want = large().get(key, {})
    if "truss_hist" not in want:
        pytest.skip(f"tests/golden/rmat_large.json has no truss summary of {key} yet (make_large.py)")
    return want


def checksum(a):
    w = (np.arange(a.size, dtype=np.uint64) % np.uint64(1000003)) + np.uint64(1)
    return int(np.sum(a.astype(np.uint64) * w, dtype=np.uint64))


def bc_sources(A, nsrc=1024, seed=1):
    """The BC workload's sources: distinct vertices with a neighbour, drawn
    with numpy's PCG64 from the seed, sorted (as make_large.py draws them)."""
    rng = np.random.default_rng(seed)
    cand = np.nonzero(np.diff(A.Rst) > 0)[0]
    return np.sort(rng.choice(cand, size=nsrc, replace=False)).astype(np.int32)


def test_bc_rmat20_4x256_fp64(gpu):
    say("bc rmat20 4x256")
    z = np.load(GOLDEN / "bc_rmat20.npz", allow_pickle=False)
    G = HostGraph.rmat(20, 16, seed=1, **RMAT)
    assert G.edges == int(z["edges"])
    src = bc_sources(G.A)
    assert np.array_equal(src, z["sources"])
    dA = gpu.upload(G.A)
    out = gpu.alloc(8 * G.n)
    gpu.bc_batch_dev(dA, src, out)
    bc = out.to_host(np.float64, G.n)
    out.free()
    dA.free()
    assert gpu.last_timing()["iterations"] > 4 * 2  # several levels per batch of 256
    vs, want = z["vertices"], z["values"]
    err = np.abs(bc[vs] - want)
    rel = err / np.maximum(np.abs(want), 1.0)
    say(f"bc s20: max rel err {rel.max():.3e} at {vs.size} vertices; sum {bc.sum():.17g} vs {float(z['total']):.17g}")
    assert np.all(rel <= BC_RTOL)
    assert abs(bc.sum() - float(z["total"])) <= BC_RTOL * abs(float(z["total"]))
    assert int(np.argmax(bc)) == int(z["argmax"])
    assert abs(bc.max() - float(z["vmax"])) <= BC_RTOL * float(z["vmax"])
    assert np.all(bc >= 0.0)


def test_batch_4096_graphs_per_graph_results(gpu):
    say("batch 4096")
    z = np.load(GOLDEN / "batch4096.npz", allow_pickle=False)
    cfg = batch_config(4096, seed=1)
    U = HostGraph.batch(**cfg)
    voff = np.asarray(U.voff, np.int64)
    assert U.A.nz > (1 << 30)  # past the int32 midpoint limit of CSR bisections
    rst = np.asarray(U.A.Rst, np.int64)
    assert np.array_equal((rst[voff[1:]] - rst[voff[:-1]]) // 2, z["edges"])
    dA = gpu.upload(U.A)
    say(f"batch: union uploaded, {U.A.nz} entries")
    del U
    r = gpu.batch_tc_ktruss(dA, voff, k=4, use_graph=True, reps=2)
    dA.free()
    say(f"batch: {r['ms']:.1f} ms per step")
    assert np.array_equal(r["triangles"].astype(np.uint64), z["triangles"])
    assert np.array_equal(r["truss_edges"].astype(np.int64), z["truss4_edges"])


@pytest.fixture(scope="module")
def trained(gpu):
    """A network trained (seeded) on timing samples of three small graphs of
    the three families."""
    graphs = [HostGraph.rmat(14, 16, seed=5, **RMAT), HostGraph.grid(128), HostGraph.chung_lu(20000, 2.0, 16, seed=5)]
    rep = selector.train_from_graphs(gpu, graphs, epochs=20, seed=1)
    say("selector trained:", rep)
    yield rep
    check(lib().gk_model_clear(gpu.handle), "gk_model_clear")


def check_graph(gpu, G, want, opts, with_support=False):
    dL = gpu.upload(G.L)
    assert gpu.tc_count(dL, opts) == want["triangles"]
    dL.free()
    dA = gpu.upload(G.A)
    truss, kmax = gpu.truss_decompose(dA, opts)
    assert kmax == want["kmax"]
    if "truss_hist" in want:
        assert np.bincount(truss).tolist() == want["truss_hist"]
        assert checksum(truss) == want["truss_checksum"]
    if with_support:
        sup = gpu.ktruss_support(G.A, opts)
        assert int(sup.sum(dtype=np.int64)) == 6 * want["triangles"]
        assert np.all(truss.astype(np.int64) <= sup.astype(np.int64) + 2)
    dA.free()
    return truss


@pytest.mark.parametrize("sel", ["learned", "heuristic"])
def test_grid2048_closed_forms(gpu, trained, sel):
    say("grid 2048", sel)
    R = 2048
    G = HostGraph.grid(R)
    assert G.edges == 3 * R * R - 4 * R + 1 == 12574721
    want = {"triangles": 2 * (R - 1) ** 2, "kmax": 3}
    assert want["triangles"] == 8380418 == large()["grid2048"]["triangles"]
    opts = LEARNED if sel == "learned" else None
    truss = check_graph(gpu, G, want, opts)
    assert np.all(truss == 3)
    sup = gpu.ktruss_support(G.A, opts)
    assert int(sup.sum(dtype=np.int64)) == 2 * 25141254  # both directions of every edge
    # 4(R-1) perimeter edges with support 1, the rest 2 (both directions)
    assert int(np.count_nonzero(sup == 1)) == 2 * 4 * (R - 1)
    assert int(np.count_nonzero(sup == 2)) == 2 * (G.edges - 4 * (R - 1))


@pytest.mark.parametrize("sel", ["learned", "heuristic"])
def test_chung_lu_4M(gpu, trained, sel):
    say("chung-lu 4M", sel)
    want = summary_of("chung_lu_4M_exp2.0_d16")
    G = HostGraph.chung_lu(1 << 22, 2.0, 16.0, seed=1)
    assert G.edges == want["edges"]
    check_graph(gpu, G, want, LEARNED if sel == "learned" else None, with_support=(sel == "learned"))


@pytest.mark.parametrize("sel", ["learned", "heuristic"])
def test_rmat22(gpu, trained, sel):
    say("rmat22", sel)
    want = summary_of("rmat22")
    G = HostGraph.rmat(22, 16, seed=1, **RMAT)
    assert G.edges == want["edges"]
    check_graph(gpu, G, want, LEARNED if sel == "learned" else None)
