"""Pins the CPU oracle (oracle/oracle.cpp) before it is trusted as the
checker: golden fixtures written by networkx/scipy, closed forms, brute force
on tiny graphs, and cross-checks between the oracle's independent
restatements.  Each check is chosen so a plausible slip in the restatement
(a dropped term, a wrong index, a transposed operand, a missing mask) fails.
"""
from math import comb

import numpy as np
import pytest
import scipy.sparse

from conftest import GOLDEN_NAMES, load_golden
from graphkokkos_amd import HostGraph, SpMtx
from oracle import oracle as O


def complete(n):
    u, v = zip(*[(i, j) for i in range(n) for j in range(i)])
    return HostGraph.from_edges(n, u, v)


def random_spmtx(rng, m, n, density, lo=-3, hi=4, sorted_unique=True):
    rows, cols, vals = [], [], []
    rst = [0]
    for i in range(m):
        k = rng.binomial(n, density)
        c = np.sort(rng.choice(n, size=k, replace=False)) if k else np.zeros(0, int)
        cols.append(c)
        vals.append(rng.integers(lo, hi, size=k))
        rst.append(rst[-1] + k)
    col = np.concatenate(cols) if cols else np.zeros(0)
    val = np.concatenate(vals) if vals else np.zeros(0)
    return SpMtx(m, n, np.array(rst, np.int32), col.astype(np.int32), val.astype(np.int32))


# ---- golden fixtures (networkx / scipy) -------------------------------------

@pytest.mark.parametrize("name", GOLDEN_NAMES)
def test_oracle_matches_golden(name):
    A, z = load_golden(name)
    G = HostGraph.from_spmtx(A)
    assert np.array_equal(G.A.Rst, A.Rst) and np.array_equal(G.A.Col, A.Col)
    for acc in (O.MSA, O.MCA):
        for par in (True, False):
            assert O.tc_sandia(G.L, acc, par) == int(z["triangles"])
            assert np.array_equal(O.support(A, acc, par), z["support"])
    tr, kmax, _ = O.truss_decompose(A)
    assert np.array_equal(tr, z["truss"])
    tp, kp = O.truss_peel(A)
    assert np.array_equal(tp, z["truss"]) and kp == kmax
    bc = O.bc_batch(A, z["sources"])
    np.testing.assert_allclose(bc, z["bc"], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("name", GOLDEN_NAMES)
def test_vertex_triangles_from_supports(name):
    """Triangles through a vertex = half the supports of its edges = diag(A^3)/2:
    the expectation the per-vertex driver is checked against."""
    A, z = load_golden(name)
    rows = np.repeat(np.arange(A.m), np.diff(A.Rst))
    half = np.bincount(rows, weights=O.support(A), minlength=A.m).astype(np.int64) // 2
    D = np.zeros((A.m, A.n), np.int64)
    D[rows, A.Col] = 1
    assert np.array_equal(half, np.diag(D @ D @ D) // 2)
    assert int(half.sum()) == 3 * int(z["triangles"])


# ---- closed forms -----------------------------------------------------------

@pytest.mark.parametrize("n", [3, 5, 8, 13])
def test_complete_graph_closed_forms(n):
    G = complete(n)
    assert O.tc_sandia(G.L) == comb(n, 3)
    assert set(O.support(G.A).tolist()) == {n - 2}
    tr, kmax, _ = O.truss_decompose(G.A)
    assert kmax == n and set(tr.tolist()) == {n}


@pytest.mark.parametrize("r", [2, 6, 9])
def test_triangulated_grid_closed_forms(r):
    G = HostGraph.grid(r)
    assert G.edges == 3 * r * r - 4 * r + 1
    assert O.tc_sandia(G.L) == 2 * (r - 1) ** 2
    s = O.support(G.A)
    # 4(r-1) perimeter edges with support 1 (counted twice in A), the rest 2
    assert np.count_nonzero(s == 1) == 2 * 4 * (r - 1)
    assert np.count_nonzero(s == 2) == 2 * (G.edges - 4 * (r - 1))
    tr, kmax, _ = O.truss_decompose(G.A)
    assert kmax == 3 and set(tr.tolist()) == {3}
    keep, _ = O.ktruss(G.A, 4)
    assert not keep.any()  # empty 4-truss


def grid_ktruss4_iterations(r: int) -> int:
    """The reference loop's iteration count (src/main.cpp:104-131, the value
    it prints as "Iterations:") for k = 4 on the r x r triangulated grid.

    Every interior edge has support 2 and every perimeter edge support 1, so
    iteration 1 removes the perimeter.  A triangle is gone from the first
    product after any of its edges is removed, and an interior edge (two
    triangles) is removed by the product after the first of its triangles
    is gone.  So a triangle d steps away from a perimeter triangle in the
    dual graph (triangles adjacent when they share an edge) is gone after
    iteration d + 1, and the last edges to go are those whose two triangles
    are both deepest.  With the diagonal of cell (i, j) running from (i, j)
    to (i+1, j+1), the upper triangle of the cell is min(2i, 2(r-2-j),
    2(r-2-i)+1, 2j+1) steps deep and the lower one min(2(r-2-i), 2j, 2i+1,
    2(r-2-j)+1); at the centre that is r - 2 for even r and r - 3 for odd r
    (where no cell holds two triangles of depth r - 2).  The deepest edge
    goes in iteration depth + 2, and the loop needs one more product that
    removes nothing: 2*floor(r/2) + 1 iterations."""
    return 2 * (r // 2) + 1


@pytest.mark.parametrize("r", range(2, 14))
def test_ktruss_iteration_count_closed_forms(r):
    G = HostGraph.grid(r)
    keep, it = O.ktruss(G.A, 4)
    assert not keep.any() and it == grid_ktruss4_iterations(r)
    # k = 3: every grid edge lies in a triangle, so the first product removes
    # nothing; k = 5: every support is below 3, so all edges go at once and
    # the second product (on the empty pattern) removes nothing
    keep, it = O.ktruss(G.A, 3)
    assert keep.all() and it == 1
    keep, it = O.ktruss(G.A, 5)
    assert not keep.any() and it == 2


@pytest.mark.parametrize("n", [4, 7])
def test_ktruss_iteration_count_pendant_and_complete(n):
    # K_n: supports n-2, so k <= n keeps everything in one iteration and
    # k = n + 1 removes everything in the first
    G = complete(n)
    for k, want_keep, want_it in ((3, True, 1), (n, True, 1), (n + 1, False, 2)):
        keep, it = O.ktruss(G.A, k)
        assert bool(keep.all() if want_keep else not keep.any()) and it == want_it, (k, it)
    # a pendant edge lies in no triangle: its removal lowers no support, so
    # k = 3 takes exactly two iterations whatever else the graph holds
    u, v = zip(*[(i, j) for i in range(n) for j in range(i)] + [(0, n)])
    G = HostGraph.from_edges(n + 1, u, v)
    keep, it = O.ktruss(G.A, 3)
    assert it == 2 and int(keep.sum()) == n * (n - 1)


def test_bipartite_has_no_triangles():
    u, v = zip(*[(i, 10 + j) for i in range(10) for j in range(7)])
    G = HostGraph.from_edges(17, u, v)
    assert O.tc_sandia(G.L) == 0
    assert not O.support(G.A).any()
    tr, kmax, _ = O.truss_decompose(G.A)
    assert kmax == 2 and set(tr.tolist()) == {2}


def test_path_and_star_bc():
    # path 0-1-...-8, all sources: bc(v) = 2 * v * (n-1-v)
    n = 9
    G = HostGraph.from_edges(n, list(range(n - 1)), list(range(1, n)))
    bc = O.bc_batch(G.A, np.arange(n))
    assert np.allclose(bc, [2 * v * (n - 1 - v) for v in range(n)])
    # star with centre 0 and 7 leaves, all sources: centre 7*6, leaves 0
    G = HostGraph.from_edges(8, [0] * 7, list(range(1, 8)))
    bc = O.bc_batch(G.A, np.arange(8))
    assert bc[0] == 42 and not bc[1:].any()


def test_bc_repeated_edge_counts_per_copy():
    # Path 0 - 1 - 2 with the edge 0-1 listed twice, source 0.  sigma(0,1) = 2
    # and sigma(0,2) = 2 (each copy is a path), so delta_0(1) = 2/2 * (1 + 0) = 1.
    # The reference's first frontier holds (1, source 0) twice and
    # bc_backward_step adds (A*W)*Numsp once per copy
    # (include/betweenness_centrality.h:236-239, 447), so vertex 1 gets 2.
    A = SpMtx(3, 3, np.array([0, 2, 5, 6], np.int32), np.array([1, 1, 0, 0, 2, 1], np.int32))
    assert O.bc_batch(A, [0]).tolist() == [0.0, 2.0, 0.0]
    # textbook Brandes on the multigraph counts the dependency once
    assert O.brandes_bc(A, [0]).tolist() == [0.0, 1.0, 0.0]
    # from source 2 the repeated edge is one hop further out: no per-copy
    # effect, sigma(2,0) = 2 and delta_2(1) = 1
    assert O.bc_batch(A, [2]).tolist() == [0.0, 1.0, 0.0]


# ---- brute force ------------------------------------------------------------

@pytest.mark.parametrize("seed", range(4))
def test_masked_product_brute_force(seed):
    rng = np.random.default_rng(seed)
    m, k, n = 23, 17, 19
    A = random_spmtx(rng, m, k, 0.3)
    B = random_spmtx(rng, k, n, 0.3)
    M = random_spmtx(rng, m, n, 0.4)
    want = O.dense_masked_product(A, B, M)
    for acc in (O.MSA, O.MCA):
        for par in (True, False):
            assert np.array_equal(O.mxmm(A, B, M, acc, par), want)


@pytest.mark.parametrize("seed,shape,dens", [(0, (23, 17, 19), 0.3), (1, (40, 40, 40), 0.1), (2, (1, 6, 3), 0.9),
                                             (3, (30, 8, 50), 0.25)])
def test_complement_mask_product_brute_force(seed, shape, dens):
    rng = np.random.default_rng(seed)
    m, k, n = shape
    A = random_spmtx(rng, m, k, dens, lo=-3, hi=4)   # explicit zeros among the values
    B = random_spmtx(rng, k, n, dens, lo=-3, hi=4)
    M = random_spmtx(rng, m, n, 0.3)
    wr, wc, wv = O.dense_cmask_product(A, B, M)
    for par in (False, True):
        rst, col, val = O.mxmm_cmask(A, B, M, parallel=par)
        assert np.array_equal(rst, wr) and np.array_equal(col, wc) and np.array_equal(val, wv)
    # an empty mask gives the whole structural product; a full one nothing
    E = SpMtx(m, n, np.zeros(m + 1, np.int32), np.zeros(0, np.int32))
    rst, col, val = O.mxmm_cmask(A, B, E)
    wr, wc, wv = O.dense_cmask_product(A, B, E)
    assert np.array_equal(rst, wr) and np.array_equal(col, wc) and np.array_equal(val, wv)
    F = SpMtx(m, n, (np.arange(m + 1) * n).astype(np.int32), np.tile(np.arange(n), m).astype(np.int32))
    assert O.mxmm_cmask(A, B, F)[0][-1] == 0


def test_complement_mask_two_hop_identity():
    """C<not A> = A*A on a symmetric pattern: its values count common
    neighbours of non-adjacent pairs (the diagonal included), so they sum to
    sum(deg^2) minus 6 per triangle; its pattern is the 2-hop pairs minus the
    edges."""
    g = HostGraph.rmat(11, 16, seed=3)
    A = g.A
    Av = SpMtx(A.m, A.n, A.Rst, A.Col, np.ones(A.nz, np.int32))
    rst, col, val = O.mxmm_cmask(Av, Av, A)
    for x, y in zip((rst, col, val), O.mxmm_cmask(Av, Av, A, parallel=False)):
        assert np.array_equal(x, y)
    deg = np.diff(A.Rst).astype(np.int64)
    assert int(val.astype(np.int64).sum()) == int((deg * deg).sum()) - 6 * O.tc_sandia(g.L)
    S = scipy.sparse.csr_matrix((np.ones(A.nz), A.Col, A.Rst), shape=(A.m, A.n))
    reach = ((S @ S) > 0).astype(np.int8)
    two = reach - reach.multiply(S > 0)
    two.eliminate_zeros()
    two = two.tocsr()
    two.sort_indices()
    assert np.array_equal(rst, two.indptr) and np.array_equal(col, two.indices)


def test_int32_wraparound_matches_dense():
    rng = np.random.default_rng(5)
    A = random_spmtx(rng, 9, 9, 0.8, lo=2**20, hi=2**21)
    B = random_spmtx(rng, 9, 9, 0.8, lo=2**20, hi=2**21)
    M = random_spmtx(rng, 9, 9, 0.9, lo=1, hi=2)
    assert np.array_equal(O.mxmm(A, B, M, O.MCA), O.dense_masked_product(A, B, M))


def test_reference_sequential_msa_stale_read():
    """_mxmm_msa_sequential returns a previous row's accumulator value at a
    mask column that gets no contribution (include/matrix_la.h:726-729);
    the parallel MSA and the MCA return 0 there."""
    # row 0: C(0,1) = A(0,0)*B(0,1) = 5; row 1: mask has column 1, no product.
    A = SpMtx(2, 1, np.array([0, 1, 1], np.int32), np.array([0], np.int32), np.array([5], np.int32))
    B = SpMtx(1, 2, np.array([0, 1], np.int32), np.array([1], np.int32), np.array([1], np.int32))
    M = SpMtx(2, 2, np.array([0, 1, 2], np.int32), np.array([1, 1], np.int32))
    assert O.mxmm(A, B, M, O.MSA, True).tolist() == [5, 0]
    assert O.mxmm(A, B, M, O.MCA, False).tolist() == [5, 0]
    assert O.mxmm(A, B, M, O.MSA, False).tolist() == [5, 0]
    assert O.mxmm(A, B, M, O.MSA_STALE, False).tolist() == [5, 5]


@pytest.mark.parametrize("seed", range(3))
def test_truss_and_bc_brute_force(seed):
    G = HostGraph.er(24, 7.0, seed=seed)
    want = O.brute_truss(G.A)
    assert np.array_equal(O.truss_decompose(G.A)[0], want)
    assert np.array_equal(O.truss_peel(G.A)[0], want)
    src = np.arange(0, 24, 3)
    np.testing.assert_allclose(O.bc_batch(G.A, src), O.brandes_bc(G.A, src), rtol=1e-12, atol=1e-12)


def test_bc_directed_brute_force():
    rng = np.random.default_rng(9)
    A = random_spmtx(rng, 30, 30, 0.12, lo=1, hi=2)
    # drop self loops, keep the pattern
    keep = np.repeat(np.arange(30), np.diff(A.Rst)) != A.Col
    rows = np.repeat(np.arange(30), np.diff(A.Rst))[keep]
    rst = np.zeros(31, np.int32)
    np.cumsum(np.bincount(rows, minlength=30), out=rst[1:])
    D = SpMtx(30, 30, rst, A.Col[keep])
    src = np.arange(0, 30, 2)
    np.testing.assert_allclose(O.bc_batch(D, src), O.brandes_bc(D, src), rtol=1e-12, atol=1e-12)
    # batches of 4 columns give the same sums as one batch
    np.testing.assert_allclose(O.bc_batch(D, src, batch=4), O.bc_batch(D, src), rtol=1e-12, atol=1e-12)


def test_bc_fp32_mode_close_to_fp64():
    G = HostGraph.rmat(9, 8, seed=2)
    src = np.arange(0, 64)
    b64 = O.bc_batch(G.A, src, fp64=True)
    b32 = O.bc_batch(G.A, src, fp64=False)
    np.testing.assert_allclose(b32, b64, rtol=2e-3, atol=1e-2)


# ---- invariants at moderate size -------------------------------------------

def test_invariants_rmat12():
    G = HostGraph.rmat(12, 16, seed=4)
    T = O.tc_sandia(G.L)
    s = O.support(G.A)
    assert int(s.sum()) == 6 * T  # every triangle adds 1 to 3 edges, each stored twice
    tr, kmax, _ = O.truss_decompose(G.A)
    assert np.all(tr <= s.astype(np.int64) + 2)
    tp, kp = O.truss_peel(G.A)
    assert np.array_equal(tr, tp) and kp == kmax
    # BC: sum_v delta_s(v) = sum over reachable t != s of (dist(s,t) - 1)
    import scipy.sparse.csgraph as cg
    src = np.array([0, 5, 77, 1000], np.int32)
    bc = O.bc_batch(G.A, src)
    d = cg.shortest_path(G.A.to_scipy(), unweighted=True, indices=src)
    fin = np.isfinite(d) & (d > 0)
    assert np.isclose(bc.sum(), (d[fin] - 1).sum())


def test_relabel_invariance_of_triangles():
    G1 = HostGraph.rmat(11, 8, seed=6, relabel=True)
    G2 = HostGraph.rmat(11, 8, seed=6, relabel=False)
    assert G1.edges == G2.edges
    assert O.tc_sandia(G1.L) == O.tc_sandia(G2.L)
