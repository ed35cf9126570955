"""Error paths of the C ABI: malformed inputs are rejected on the host,
before any kernel runs, with GK_EINVAL and a message naming the problem."""
import numpy as np
import pytest

from graphkokkos_amd import GraphHipError, HostGraph, SpMtx

pytestmark = pytest.mark.gpu


def csr(rows, n=None):
    n = n if n is not None else len(rows)
    rst = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    col = np.array([c for r in rows for c in r], np.int32)
    return SpMtx(len(rows), n, rst, col, np.ones(len(col), np.int32))


def test_column_out_of_range(gpu):
    A = csr([[1], [0, 3], [1]], n=3)            # column 3 in a 3-column matrix
    with pytest.raises(GraphHipError, match="outside"):
        gpu.masked_spgemm(A, A, A)


def test_decreasing_row_pointers(gpu):
    A = SpMtx(3, 3, np.array([0, 2, 1, 2], np.int32), np.array([1, 2], np.int32), np.ones(2, np.int32))
    with pytest.raises(GraphHipError, match="row pointers"):
        gpu.masked_spgemm(A, A, A)


def test_unsorted_mask_rejected_by_c_abi(gpu):
    A = csr([[1, 2], [2, 0], [0, 1]])           # row 1 unsorted
    with pytest.raises(GraphHipError, match="strictly increasing"):
        gpu.masked_spgemm(A, A, A)
    # the reference-named entry point accepts it (mask canonicalised on the host)
    from graphkokkos_amd import mxmm_msa
    C = SpMtx()
    mxmm_msa(True, A, A, A, C)
    assert C.Val.tolist() == [1, 1, 1, 1, 1, 1]


def test_truss_needs_symmetric_pattern(gpu):
    A = csr([[1], [], []])                      # one directed edge
    with pytest.raises(GraphHipError, match="symmetric"):
        gpu.truss_decompose(A)
    S = csr([[0, 1], [0], []])                  # self loop on vertex 0
    with pytest.raises(GraphHipError, match="symmetric"):
        gpu.ktruss_support(S)


def test_valid_inputs_still_accepted(gpu):
    G = HostGraph.grid(6)
    tr, kmax = gpu.truss_decompose(G.A)
    assert kmax == 3 and np.all(tr == 3)
