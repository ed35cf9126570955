"""GPU parity: libgraphhip.so (through its C ABI) against the oracle on the
same seeded inputs.  Integer outputs (triangle counts, masked-product values,
supports, k-truss patterns, truss numbers) must be bit-exact; BC (FP64) must
agree within 1e-9 relative (absolute floor 1e-9 for exact zeros).
"""
import sys

import numpy as np
import pytest

from conftest import GOLDEN_NAMES, load_golden
from graphkokkos_amd import (ACC_DENSE_GLOBAL, ACC_HASH_GLOBAL, ACC_MCA_WAVE, ACC_MSA_LDS, SEL_FORCED,
                             HostGraph, SpMtx, betweenness_centrality_batch, k_truss, mxmm_mca, mxmm_msa,
                             triangle_counting_masked_sandia)
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ACCS = [ACC_MSA_LDS, ACC_MCA_WAVE, ACC_HASH_GLOBAL, ACC_DENSE_GLOBAL]
BC_RTOL = 1e-9


def forced(acc):
    return {"selector": SEL_FORCED, "forced_accum": acc}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def assert_bc_close(got, want):
    err = np.abs(got - want)
    tol = BC_RTOL * np.maximum(np.abs(want), 1.0)
    assert np.all(err <= tol), f"max abs err {err.max():.3e}, max rel err {(err / np.maximum(np.abs(want), 1e-300)).max():.3e}"


# ---- golden fixtures ---------------------------------------------------------

@pytest.mark.parametrize("name", GOLDEN_NAMES)
def test_golden_all(gpu, name):
    say("golden", name)
    A, z = load_golden(name)
    G = HostGraph.from_spmtx(A)
    for acc in ACCS:
        assert gpu.tc_count(G.L, forced(acc)) == int(z["triangles"]), acc
    assert gpu.tc_count(G.L) == int(z["triangles"])
    for hw in (0, 32, 1024):
        assert gpu.tc_count(G.L, {**forced(ACC_MSA_LDS), "hub_window": hw}) == int(z["triangles"]), hw
    assert np.array_equal(gpu.ktruss_support(A).astype(np.int64), z["support"].astype(np.int64))
    tr, kmax = gpu.truss_decompose(A)
    assert np.array_equal(tr, z["truss"])
    bc = gpu.bc_batch_f64(A, z["sources"])
    assert_bc_close(bc, z["bc"])


# ---- triangle counting ------------------------------------------------------

@pytest.mark.parametrize("scale,ef,seed", [(8, 8, 1), (10, 16, 2), (12, 16, 3), (14, 16, 1)])
def test_tc_rmat_all_accumulators(gpu, scale, ef, seed):
    say("tc rmat", scale, ef, seed)
    G = HostGraph.rmat(scale, ef, seed=seed)
    want = O.tc_sandia(G.L)
    assert gpu.tc_count(G.L) == want
    for acc in ACCS:
        assert gpu.tc_count(G.L, forced(acc)) == want, acc


@pytest.mark.parametrize("kind", ["er", "chunglu", "grid", "grid-norelabel"])
def test_tc_other_graphs(gpu, kind):
    say("tc", kind)
    if kind == "er":
        G = HostGraph.er(5000, 24.0, seed=3)
    elif kind == "chunglu":
        G = HostGraph.chung_lu(20000, 2.0, 16.0, seed=5)
    elif kind == "grid":
        G = HostGraph.grid(64)
    else:
        G = HostGraph.grid(50, relabel=False)
    want = O.tc_sandia(G.L)
    assert gpu.tc_count(G.L) == want
    for acc in ACCS:
        assert gpu.tc_count(G.L, forced(acc)) == want, acc
    if kind.startswith("grid"):
        r = 64 if kind == "grid" else 50
        assert want == 2 * (r - 1) ** 2


def test_tc_per_row_and_drop_in(gpu):
    G = HostGraph.rmat(11, 16, seed=8)
    d = gpu.upload(G.L)
    rows = gpu.tc_rows(d)
    C = SpMtx()
    mxmm_msa(True, G.L, G.L, G.L, C)
    rows_of = np.repeat(np.arange(G.n), np.diff(G.L.Rst))
    per_row = np.bincount(rows_of, weights=C.Val, minlength=G.n).astype(np.int64)
    assert np.array_equal(rows.astype(np.int64), per_row)
    assert triangle_counting_masked_sandia(G.A, mxmm_msa, True) == O.tc_sandia(G.L)
    assert triangle_counting_masked_sandia(G.A, mxmm_mca, True) == O.tc_sandia(G.L)


# Hub bitmaps: B rows below hub_window read as bitmap words where that is
# cheaper.  Windows that are not multiples of 32, wider than the graph, and
# the largest one; 0 turns them off (every B row streamed, the clamped
# stream of the window kernels); the window kernels forced, the heuristic,
# per-row counts.
HUB_WINDOWS = [0, 37, 64, 4096, 32768]


@pytest.mark.parametrize("kind", ["rmat12", "rmat14", "chunglu", "grid", "er"])
def test_tc_hub_bitmaps(gpu, kind):
    say("tc hub", kind)
    G = {"rmat12": lambda: HostGraph.rmat(12, 16, seed=3), "rmat14": lambda: HostGraph.rmat(14, 16, seed=1),
         "chunglu": lambda: HostGraph.chung_lu(20000, 2.0, 16.0, seed=5), "grid": lambda: HostGraph.grid(64),
         "er": lambda: HostGraph.er(5000, 24.0, seed=3)}[kind]()
    want = O.tc_sandia(G.L)
    d = gpu.upload(G.L)
    rows_want = gpu.tc_rows(d)
    assert int(rows_want.sum()) == want
    for hw in HUB_WINDOWS:
        assert gpu.tc_count(d, {"hub_window": hw}) == want, hw
        for acc in ACCS:
            assert gpu.tc_count(d, {**forced(acc), "hub_window": hw}) == want, (hw, acc)
        # long rows through the block-wide kernel, short rows through the wave
        # kernel; bit tests never, at the cost model's weights, and always
        # (weight 1 picks them wherever they are fewer)
        for sm in (1, 128):
            for pr in (0, 1, 4):
                o = {**forced(ACC_MSA_LDS), "hub_window": hw, "small_max_mask": sm, "hub_probe": pr}
                assert gpu.tc_count(d, o) == want, (hw, sm, pr)
        assert np.array_equal(gpu.tc_rows(d, {"hub_window": hw}), rows_want), hw
    # split rows: hub words, bit tests and streamed entries of one row in
    # several tasks (a task's bit tests start at the row's first mask column)
    for pr in (0, 1):
        assert gpu.tc_count(d, {**forced(ACC_MSA_LDS), "hub_window": 4096, "split_flop": 64, "hub_probe": pr}) == want
    d.free()


def test_tc_edge_cases(gpu):
    # no edges, one vertex, isolated vertices only
    for n in (1, 2, 50):
        G = HostGraph.from_edges(n, [], [])
        assert gpu.tc_count(G.L) == 0
    # a single triangle among isolated vertices
    G = HostGraph.from_edges(10, [7, 8, 8], [3, 3, 7])
    assert gpu.tc_count(G.L) == 1


# ---- general masked product -------------------------------------------------

def rand_csr(rng, m, n, density, lo=-5, hi=6):
    rst, cols, vals = [0], [], []
    for _ in range(m):
        k = rng.binomial(n, density)
        cols.append(np.sort(rng.choice(n, size=k, replace=False)))
        vals.append(rng.integers(lo, hi, size=k))
        rst.append(rst[-1] + k)
    return SpMtx(m, n, np.array(rst, np.int32), np.concatenate(cols).astype(np.int32),
                 np.concatenate(vals).astype(np.int32))


@pytest.mark.parametrize("seed,shape", [(0, (37, 29, 41)), (1, (200, 150, 300)), (2, (1, 5, 1)),
                                        (3, (500, 500, 500))])
def test_masked_spgemm_values(gpu, seed, shape):
    say("spgemm", seed, shape)
    rng = np.random.default_rng(seed)
    m, k, n = shape
    A = rand_csr(rng, m, k, 0.2)
    B = rand_csr(rng, k, n, 0.2)
    M = rand_csr(rng, m, n, 0.3, 1, 2)
    want = O.mxmm(A, B, M, O.MCA)
    assert np.array_equal(want, O.dense_masked_product(A, B, M))
    assert np.array_equal(gpu.masked_spgemm(A, B, M), want)
    for acc in ACCS:
        assert np.array_equal(gpu.masked_spgemm(A, B, M, forced(acc)), want), acc


def test_masked_spgemm_split_rows_and_wrap(gpu):
    """Long A rows are cut into several tasks (atomic merge); large values
    wrap around int32 as the reference's int arithmetic does."""
    rng = np.random.default_rng(4)
    A = rand_csr(rng, 8, 3000, 0.5, 2**16, 2**17)
    B = rand_csr(rng, 3000, 64, 0.3, 2**15, 2**16)
    M = rand_csr(rng, 8, 64, 0.7, 1, 2)
    want = O.mxmm(A, B, M, O.MSA)
    for acc in ACCS:
        got = gpu.masked_spgemm(A, B, M, {**forced(acc), "split_flop": 4096})
        assert np.array_equal(got, want), acc


def test_mxmm_drop_in_semantics(gpu):
    rng = np.random.default_rng(6)
    A = rand_csr(rng, 40, 30, 0.2)
    B = rand_csr(rng, 30, 50, 0.2)
    M = rand_csr(rng, 40, 50, 0.3, 1, 2)
    C = SpMtx()
    mxmm_msa(True, A, B, M, C)
    assert C.m == M.m and C.n == M.n and C.nz == M.nz
    assert np.array_equal(C.Rst, M.Rst) and np.array_equal(C.Col, M.Col)
    assert np.array_equal(C.Val, O.mxmm(A, B, M))
    with pytest.raises(Exception):
        gpu.masked_spgemm(A, A, M)  # A.n != B.m


# ---- supports and k-truss ---------------------------------------------------

@pytest.mark.parametrize("scale", [10, 13])
def test_support_rmat(gpu, scale):
    G = HostGraph.rmat(scale, 16, seed=scale)
    want = O.support(G.A)
    got = gpu.ktruss_support(G.A)
    assert np.array_equal(got.astype(np.int64), want.astype(np.int64))


@pytest.mark.parametrize("kind", ["rmat12", "star_paths", "er"])
def test_support_window_kernel_segments(gpu, kind):
    """Supports with every row in the block-wide window kernel, whose waves
    each stream a contiguous segment of a task's B rows and must report the
    A entry of every hit: whole rows, rows split into tasks of a few
    entries, and a graph whose hub rows meet many empty and one-entry B rows
    (segments that start among them, and streams shorter than a wave)."""
    say("support window segments", kind)
    if kind == "star_paths":
        n = 3000
        rng = np.random.default_rng(7)
        hubs = np.arange(8)
        edges = [(h, v) for h in hubs for v in range(8, n) if rng.random() < 0.6]
        edges += [(v, v + 1) for v in range(8, n - 1) if rng.random() < 0.5]
        e = np.array(edges, np.int64)
        G = HostGraph.from_edges(n, e[:, 0], e[:, 1])
    else:
        G = {"rmat12": lambda: HostGraph.rmat(12, 16, seed=4), "er": lambda: HostGraph.er(4000, 20.0, seed=9)}[kind]()
    want = O.support(G.A).astype(np.int64)
    tc_want = O.tc_sandia(G.L)
    d = gpu.upload(G.L)
    for split in (1 << 18, 64, 5):
        o = {**forced(ACC_MSA_LDS), "small_max_mask": 1, "split_flop": split}
        got = gpu.ktruss_support(G.A, o).astype(np.int64)
        assert np.array_equal(got, want), (kind, split)
        # triangle counting through the same kernel with hub bitmaps (the
        # default) and without them (the clamped stream that the supports use)
        for hw in (32768, 0):
            assert gpu.tc_count(d, {**o, "hub_window": hw}) == tc_want, (kind, split, hw)
    d.free()


@pytest.mark.parametrize("k", [3, 4, 6, 10])
def test_ktruss_single_k(gpu, k):
    G = HostGraph.rmat(11, 16, seed=7)
    keep, it = O.ktruss(G.A, k)
    res, it_gpu = gpu.ktruss(G.A, k)
    rows = np.repeat(np.arange(G.n), np.diff(G.A.Rst))
    want_rst = np.zeros(G.n + 1, np.int64)
    np.cumsum(np.bincount(rows[keep], minlength=G.n), out=want_rst[1:])
    assert np.array_equal(res.Rst.astype(np.int64), want_rst)
    assert np.array_equal(res.Col, G.A.Col[keep])
    assert it_gpu == it
    assert k_truss(G.A, k, mxmm_msa, True) == res


@pytest.mark.parametrize("r", [2, 3, 8, 13, 40])
def test_ktruss_iteration_count_closed_form(gpu, r):
    """gk_ktruss's iteration count (the reference's "Iterations:" line,
    src/main.cpp:104-138) on the triangulated r x r grid: 2*floor(r/2)+1
    for k = 4 (one perimeter layer peeled per product; derivation in
    test_oracle.grid_ktruss4_iterations), 1 for k = 3, 2 for k = 5."""
    say("ktruss iterations grid", r)
    G = HostGraph.grid(r)
    for k, want_it, want_nz in ((4, 2 * (r // 2) + 1, 0), (3, 1, G.A.nz), (5, 2, 0)):
        res, it = gpu.ktruss(G.A, k)
        assert it == want_it and res.nz == want_nz, (k, it, res.nz)


@pytest.mark.parametrize("scale,seed", [(10, 1), (12, 2), (14, 1), (16, 1)])
def test_truss_decomposition_rmat(gpu, scale, seed):
    say("truss", scale)
    G = HostGraph.rmat(scale, 16, seed=seed)
    want, kmax = O.truss_peel(G.A)
    got, kg = gpu.truss_decompose(G.A)
    assert kg == kmax
    assert np.array_equal(got, want)


def test_truss_other_graphs(gpu):
    for G in (HostGraph.grid(40), HostGraph.er(3000, 20.0, seed=1), HostGraph.chung_lu(5000, 2.1, 16.0, seed=2)):
        want, kmax = O.truss_peel(G.A)
        got, kg = gpu.truss_decompose(G.A)
        assert kg == kmax and np.array_equal(got, want)


def test_truss_edge_cases(gpu):
    G = HostGraph.from_edges(5, [], [])
    got, kg = gpu.truss_decompose(G.A)
    assert got.size == 0 and kg == 0
    G = HostGraph.from_edges(6, [1, 2, 3], [0, 1, 2])  # a path: all truss 2
    got, kg = gpu.truss_decompose(G.A)
    assert kg == 2 and set(got.tolist()) == {2}


# ---- betweenness ------------------------------------------------------------

@pytest.mark.parametrize("scale,nsrc", [(8, 40), (11, 256), (13, 300)])
def test_bc_rmat(gpu, scale, nsrc):
    say("bc", scale, nsrc)
    G = HostGraph.rmat(scale, 16, seed=3)
    rng = np.random.default_rng(scale)
    src = rng.choice(G.n, size=nsrc, replace=False).astype(np.int32)
    want = O.bc_batch(G.A, src)
    got = gpu.bc_batch_f64(G.A, src)
    assert_bc_close(got, want)


def test_bc_directed_and_reference_api(gpu):
    rng = np.random.default_rng(12)
    n = 300
    rst, cols = [0], []
    for i in range(n):
        c = np.sort(rng.choice(np.setdiff1d(np.arange(n), [i]), size=rng.integers(0, 6), replace=False))
        cols.append(c)
        rst.append(rst[-1] + len(c))
    D = SpMtx(n, n, np.array(rst, np.int32), np.concatenate(cols).astype(np.int32))
    src = np.arange(0, n, 3, dtype=np.int32)
    assert_bc_close(gpu.bc_batch_f64(D, src), O.brandes_bc(D, src))
    # reference signature: sources 0..batch-1, float32 result
    G = HostGraph.rmat(9, 16, seed=5)
    got = betweenness_centrality_batch(True, G.A, 64)
    assert got.dtype == np.float32
    want = O.bc_batch(G.A, np.arange(64))
    np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-3)


def test_bc_edge_cases(gpu):
    G = HostGraph.from_edges(4, [], [])
    assert not gpu.bc_batch_f64(G.A, [0, 1, 2, 3]).any()
    G = HostGraph.from_edges(9, list(range(8)), list(range(1, 9)))  # path
    got = gpu.bc_batch_f64(G.A, np.arange(9))
    assert np.allclose(got, [2 * v * (8 - v) for v in range(9)], rtol=1e-12)


def test_bc_repeated_edges(gpu):
    say("bc repeated edges")
    # path 0 - 1 - 2 with the edge 0-1 listed twice (see the oracle's test):
    # the reference counts vertex 1's dependency on source 0 once per copy
    A = SpMtx(3, 3, np.array([0, 2, 5, 6], np.int32), np.array([1, 1, 0, 0, 2, 1], np.int32))
    assert gpu.bc_batch_f64(A, [0]).tolist() == [0.0, 2.0, 0.0]
    assert gpu.bc_batch_f64(A, [2]).tolist() == [0.0, 1.0, 0.0]
    # a larger pattern with unsorted rows and ~10% repeated entries, two
    # batches of sources
    rng = np.random.default_rng(7)
    G = HostGraph.rmat(10, 8, seed=7)
    rows = np.repeat(np.arange(G.n), np.diff(G.A.Rst))
    e = np.stack([rows, G.A.Col], 1)
    e = np.concatenate([e, e[rng.choice(len(e), len(e) // 10, replace=False)]])
    e = e[rng.permutation(len(e))]
    e = e[np.argsort(e[:, 0], kind="stable")]
    rst = np.concatenate([[0], np.cumsum(np.bincount(e[:, 0], minlength=G.n))]).astype(np.int32)
    D = SpMtx(G.n, G.n, rst, e[:, 1].astype(np.int32))
    src = rng.choice(np.flatnonzero(np.diff(rst) > 0), size=300, replace=False).astype(np.int32)
    assert_bc_close(gpu.bc_batch_f64(D, src), O.bc_batch(D, src))


def hub_batches_graph():
    """Two heavy vertices (more neighbours than the one-wave kernels take):
    A = 0 with leaves 1..2600 (several chunks), B = 2601 with leaves 1..256
    and 2602..3445 (one chunk); far vertex 3702 + j hangs off leaf j + 1.
    Sources: the 256 far vertices (A and B on level 2 for every one), then
    leaves 1..256 (A and B on level 1 for every one).  In the second batch
    every source has reached both hubs before level 2, and their level-2
    frontier words must not keep the first batch's bits."""
    u = list(range(1, 2601)) + list(range(1, 257)) + list(range(2602, 3446)) + [3702 + j for j in range(256)]
    v = [0] * 2600 + [2601] * (256 + 844) + [j + 1 for j in range(256)]
    G = HostGraph.from_edges(3958, u, v)
    src = np.array([3702 + j for j in range(256)] + list(range(1, 257)), np.int32)
    return G, src


def test_bc_hubs_reached_earlier_in_a_later_batch(gpu):
    say("bc hub levels across batches")
    G, src = hub_batches_graph()
    assert np.diff(G.A.Rst)[0] == 2600 and np.diff(G.A.Rst)[2601] == 1100
    want = O.bc_batch(G.A, src)
    assert_bc_close(gpu.bc_batch_f64(G.A, src), want)
    # the batches in the other order, and each alone
    assert_bc_close(gpu.bc_batch_f64(G.A, src[::-1].copy()), want)
    assert_bc_close(gpu.bc_batch_f64(G.A, src[:256]) + gpu.bc_batch_f64(G.A, src[256:]), want)
