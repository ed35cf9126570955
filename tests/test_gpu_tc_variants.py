"""The reference's two unwired triangle-counting drivers (src/main.cpp:16-66):
triangle_counting_vertex (row sums of SQ<A> = A*A, halved, in int
arithmetic) and triangle_counting_masked_lu (sum of C<A> = L*U, halved).
Per-vertex counts are checked against the golden networkx supports (the
triangles through a vertex are half the supports of its edges), against
diag(A^3)/2 by brute force, and against the oracle's product; the fused GPU
path against the reference's loop around the product."""
import sys

import numpy as np
import pytest

from conftest import GOLDEN_NAMES, load_golden
from graphkokkos_amd import (HostGraph, SpMtx, mxmm_msa, triangle_counting_masked_lu,
                             triangle_counting_vertex)
from graphkokkos_amd._lib import GraphHipError
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def wrapped(isParallel, A, B, M, C_):
    """Any product that is not the library's: the drivers take their loop."""
    mxmm_msa(isParallel, A, B, M, C_)


def halved_row_sums(A: SpMtx, vals) -> np.ndarray:
    rows = np.repeat(np.arange(A.m), np.diff(A.Rst))
    s = np.zeros(A.m, np.int64)
    np.add.at(s, rows, np.asarray(vals, np.int64))
    return ((s & 0xFFFFFFFF).astype(np.uint32).view(np.int32) >> 1).astype(np.int32)


@pytest.mark.parametrize("name", GOLDEN_NAMES)
def test_vertex_counts_golden(gpu, name):
    A, z = load_golden(name)
    say("tc vertex", name)
    want = halved_row_sums(A, z["support"])
    D = np.zeros((A.m, A.n), np.int64)
    D[np.repeat(np.arange(A.m), np.diff(A.Rst)), A.Col] = 1
    assert np.array_equal(want, (np.diag(D @ D @ D) // 2).astype(np.int32))
    assert np.array_equal(gpu.tc_vertex(A), want)
    assert np.array_equal(triangle_counting_vertex(A, mxmm_msa), want)
    assert np.array_equal(triangle_counting_vertex(A, wrapped), want)
    assert int(want.sum()) == 3 * int(z["triangles"])
    assert triangle_counting_masked_lu(A, mxmm_msa) == int(z["triangles"])
    assert triangle_counting_masked_lu(A, wrapped) == int(z["triangles"])


def test_vertex_counts_with_values_wrap_like_int(gpu):
    rng = np.random.default_rng(4)
    n = 50
    rst, cols = [0], []
    for _ in range(n):
        c = np.sort(rng.choice(n, size=rng.integers(0, 12), replace=False))
        cols.append(c)
        rst.append(rst[-1] + len(c))
    col = np.concatenate(cols).astype(np.int32)
    val = rng.integers(1 << 28, 1 << 30, size=len(col)).astype(np.int32)  # sums wrap
    A = SpMtx(n, n, np.array(rst, np.int32), col, val, "MCIG")
    want = halved_row_sums(A, O.mxmm(A, A, A))
    assert np.array_equal(gpu.tc_vertex(A), want)
    assert np.array_equal(triangle_counting_vertex(A, wrapped), want)


def test_vertex_counts_rmat(gpu):
    g = HostGraph.rmat(16, 16, seed=1)
    A = g.A
    got = gpu.tc_vertex(A)
    assert np.array_equal(got, halved_row_sums(A, O.support(A)))
    assert int(got.astype(np.int64).sum()) == 3 * O.tc_sandia(g.L)
    assert triangle_counting_masked_lu(A, mxmm_msa) == O.tc_sandia(g.L)


def test_vertex_counts_need_sorted_square_rows(gpu):
    A = SpMtx(3, 3, np.array([0, 2, 3, 4], np.int32), np.array([2, 1, 0, 0], np.int32))
    with pytest.raises(GraphHipError):
        gpu.tc_vertex(A)
    R = SpMtx(2, 3, np.array([0, 1, 2], np.int32), np.array([1, 2], np.int32))
    with pytest.raises(GraphHipError):
        gpu.tc_vertex(R)
    E = SpMtx(0, 0, np.zeros(1, np.int32), np.zeros(0, np.int32))
    assert gpu.tc_vertex(E).shape == (0,)
    # the driver takes the loop for unsorted rows, as the reference would
    S = HostGraph.grid(4).A
    rows = np.repeat(np.arange(S.m), np.diff(S.Rst))
    perm = np.lexsort((-S.Col, rows))  # columns descending within each row
    Su = SpMtx(S.m, S.n, S.Rst, S.Col[perm])
    assert np.array_equal(triangle_counting_vertex(Su, mxmm_msa), gpu.tc_vertex(S))
