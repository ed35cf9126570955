"""Complement-mask product C<not M> = A*B (mxmm_msa_cmask,
include/matrix_la.h:733-934) on the GPU against the oracle's restatement of
the reference's two-pass MSA and against a dense brute force: rows,
columns and int32 values must agree entry for entry, whichever way the rows
go (DENSE_ROWS: -1 the library's per-row choice, 0 every row sorted, 1
every row dense, written through a pool, 2 every row dense, counted
first)."""
import sys

import numpy as np
import pytest

from graphkokkos_amd import HostGraph, SpMtx, mxmm_msa_cmask
from graphkokkos_amd._lib import GraphHipError
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DENSE_ROWS = [-1, 0, 1, 2]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rand_csr(rng, m, n, density, lo=-9, hi=10, sort=True, values=True):
    rst, cols = [0], []
    for _ in range(m):
        c = rng.choice(n, size=min(n, rng.binomial(n, density)), replace=False)
        cols.append(np.sort(c) if sort else c)
        rst.append(rst[-1] + len(c))
    col = np.concatenate(cols).astype(np.int32) if cols else np.zeros(0, np.int32)
    val = rng.integers(lo, hi, size=len(col)).astype(np.int32) if values else None
    return SpMtx(m, n, np.array(rst, np.int32), col, val, "MCIG")


def check(got: SpMtx, want):
    rst, col, val = want
    assert np.array_equal(got.Rst, rst)
    assert np.array_equal(got.Col, col)
    assert np.array_equal(np.asarray(got.Val, np.int32), np.asarray(val, np.int32))


@pytest.mark.parametrize("dense", DENSE_ROWS)
@pytest.mark.parametrize("seed,m,k,n,da,db,dm", [
    (1, 40, 30, 50, 0.1, 0.1, 0.2),
    (2, 64, 64, 64, 0.2, 0.2, 0.0),    # empty mask: the plain product's pattern
    (3, 7, 100, 3, 0.3, 0.5, 0.6),
    (4, 200, 150, 120, 0.03, 0.05, 0.1),
    (5, 1, 1, 1, 1.0, 1.0, 0.0),
    (6, 90, 1, 90, 0.5, 0.5, 0.3),     # outer products
])
def test_random_against_oracle_and_brute_force(gpu, seed, m, k, n, da, db, dm, dense):
    say("cmask random", seed, m, k, n, dense)
    rng = np.random.default_rng(seed)
    A, B, M = rand_csr(rng, m, k, da), rand_csr(rng, k, n, db), rand_csr(rng, m, n, dm, values=False)
    got = gpu.masked_spgemm_cmask(A, B, M, dense)
    assert (got.m, got.n) == (M.m, M.n)
    check(got, O.mxmm_cmask(A, B, M))
    check(got, O.dense_cmask_product(A, B, M))


@pytest.mark.parametrize("dense", DENSE_ROWS)
def test_full_mask_gives_empty_result(gpu, dense):
    rng = np.random.default_rng(7)
    A, B = rand_csr(rng, 30, 20, 0.3), rand_csr(rng, 20, 25, 0.3)
    full = SpMtx(30, 25, np.arange(0, 30 * 25 + 1, 25, dtype=np.int32), np.tile(np.arange(25), 30).astype(np.int32))
    got = gpu.masked_spgemm_cmask(A, B, full, dense)
    assert got.nz == 0 and np.array_equal(got.Rst, np.zeros(31, np.int32))


@pytest.mark.parametrize("dense", DENSE_ROWS)
def test_cancelling_terms_keep_explicit_zeros_and_wrap(gpu, dense):
    # row 0: 1*5 + 1*(-5) = 0 at column 0 (kept); row 1: 2^30 * 5 wraps to 2^30
    A = SpMtx(2, 2, np.array([0, 2, 3], np.int32), np.array([0, 1, 0], np.int32),
              np.array([1, 1, 1 << 30], np.int32))
    B = SpMtx(2, 2, np.array([0, 1, 2], np.int32), np.array([0, 0], np.int32), np.array([5, -5], np.int32))
    B2 = SpMtx(2, 2, np.array([0, 1, 2], np.int32), np.array([1, 1], np.int32), np.array([4, 3], np.int32))
    M = SpMtx(2, 2, np.array([0, 0, 0], np.int32), np.zeros(0, np.int32))
    got = gpu.masked_spgemm_cmask(A, B, M, dense)
    assert got.Rst.tolist() == [0, 1, 2] and got.Col.tolist() == [0, 0]
    assert np.asarray(got.Val).tolist() == [0, 1 << 30]
    check(gpu.masked_spgemm_cmask(A, B2, M, dense), O.dense_cmask_product(A, B2, M))


@pytest.mark.parametrize("dense", DENSE_ROWS)
def test_unsorted_operands_and_repeated_mask_columns(gpu, dense):
    rng = np.random.default_rng(11)
    A = rand_csr(rng, 50, 40, 0.15, sort=False)
    B = rand_csr(rng, 40, 60, 0.15, sort=False)
    M = rand_csr(rng, 50, 60, 0.2, sort=False, values=False)
    # repeat some mask entries in place (rows stay contiguous)
    rows = np.repeat(np.arange(M.m), np.diff(M.Rst))
    rep = rng.random(M.nz) < 0.2
    col = np.concatenate([np.concatenate([M.row(i), M.row(i)[rep[M.Rst[i]:M.Rst[i + 1]]]]) for i in range(M.m)])
    cnt = np.bincount(rows, minlength=M.m) + np.bincount(rows[rep], minlength=M.m)
    Md = SpMtx(M.m, M.n, np.concatenate([[0], np.cumsum(cnt)]).astype(np.int32), col.astype(np.int32))
    want = O.dense_cmask_product(A, B, Md)
    check(gpu.masked_spgemm_cmask(A, B, Md, dense), want)
    # the reference-named entry point fills C in place
    Cout = SpMtx()
    mxmm_msa_cmask(True, A, B, Md, Cout)
    check(Cout, want)


def test_device_entry_point_needs_a_sorted_mask(gpu):
    rng = np.random.default_rng(12)
    A = rand_csr(rng, 10, 10, 0.3)
    M = SpMtx(10, 10, np.array([0, 2] + [2] * 9, np.int32), np.array([5, 1], np.int32))
    dA, dM = gpu.upload(A), gpu.upload(M)
    with pytest.raises(GraphHipError):
        gpu.masked_spgemm_cmask_dev(dA, dA, dM)
    dB = gpu.upload(rand_csr(rng, 9, 10, 0.3))
    with pytest.raises(GraphHipError):   # A.n != B.m
        gpu.masked_spgemm_cmask_dev(dA, dB, gpu.upload(rand_csr(rng, 10, 10, 0.1, values=False)))


@pytest.mark.parametrize("dense", DENSE_ROWS)
def test_wide_products_cut_into_windows(gpu, dense):
    """C wider than one dense window (32768 columns): each B row is cut to
    the window by searches; with unsorted B rows the dense way is off."""
    rng = np.random.default_rng(21)
    n = 100_003
    A = rand_csr(rng, 40, 60, 0.3)
    B = rand_csr(rng, 60, n, 0.003)
    M = rand_csr(rng, 40, n, 0.002, values=False)
    want = O.mxmm_cmask(A, B, M)
    check(gpu.masked_spgemm_cmask(A, B, M, dense), want)
    Bu = rand_csr(rng, 60, n, 0.003, sort=False)
    check(gpu.masked_spgemm_cmask(A, Bu, M, dense), O.mxmm_cmask(A, Bu, M))


@pytest.mark.parametrize("dense", DENSE_ROWS)
def test_two_hop_pairs_of_an_rmat_graph_span_several_chunks(gpu, dense):
    """R-MAT scale 14: A*A has 156.8 M product terms, more than one sort
    chunk (2^27 terms), so the chunk concatenation is exercised when the
    rows are sorted."""
    g = HostGraph.rmat(14, 16, seed=1)
    A = g.A
    Av = SpMtx(A.m, A.n, A.Rst, A.Col, np.ones(A.nz, np.int32))
    deg = np.diff(A.Rst).astype(np.int64)
    assert int((deg * deg).sum()) > 1 << 27
    say("cmask rmat s14", A.m, A.nz, dense)
    got = gpu.masked_spgemm_cmask(Av, Av, A, dense)
    check(got, O.mxmm_cmask(Av, Av, A))
    # a size-independent check: every value counts common neighbours, so
    # the values sum to the wedges minus those closed by an edge
    assert int(np.asarray(got.Val, np.int64).sum()) == int((deg * deg).sum()) - 6 * O.tc_sandia(g.L)


def test_two_hop_pairs_rmat16_two_windows(gpu):
    """R-MAT scale 16 (65,536 columns: two dense windows), the library's
    own choice of rows, against the oracle entry for entry."""
    g = HostGraph.rmat(16, 16, seed=1)
    A = g.A
    say("cmask rmat s16", A.m, A.nz)
    got = gpu.masked_spgemm_cmask(A, A, A)
    t = gpu.last_timing()
    assert t["iterations"] > 0  # some rows took the dense way
    check(got, O.mxmm_cmask(A, A, A))


def test_result_refilled_and_grown(gpu):
    """A result handed back is refilled: smaller products reuse its storage,
    larger ones grow it; operands and uploaded CSRs are refused as C."""
    rng = np.random.default_rng(31)
    cases = [(120, 90, 150, 0.08), (30, 20, 40, 0.1), (200, 150, 300, 0.06)]
    C_ = None
    for m, k, n, dens in cases:
        A, B = rand_csr(rng, m, k, dens), rand_csr(rng, k, n, dens)
        M = rand_csr(rng, m, n, 0.05, values=False)
        dA, dB, dM = gpu.upload(A), gpu.upload(B), gpu.upload(M)
        C_ = gpu.masked_spgemm_cmask_dev(dA, dB, dM, -1, out=C_)
        check(C_.to_host(values=True), O.mxmm_cmask(A, B, M))
        with pytest.raises(GraphHipError):
            gpu.masked_spgemm_cmask_dev(dA, dB, dM, -1, out=dA)
        for d in (dA, dB, dM):
            d.free()
    C_.free()
