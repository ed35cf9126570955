"""gk_bc_batch_dev_noalloc (the allocation-free BC entry point) through the
C ABI against gk_bc_batch_dev, the entry point whose sweeps it shares: two
calls in a row, the second on a smaller graph with fewer sources (the
scratch array keeps the first call's contents), a directed graph with its
transpose, and the size and precondition errors.  The same kernels run in
the same order on the same data, so the scores must be identical; they are
also checked against the oracle at 1e-9 relative."""
import numpy as np
import pytest

from graphkokkos_amd import HostGraph, SpMtx
from graphkokkos_amd._lib import GraphHipError as GkError, lib
from oracle import oracle as O
from test_gpu_parity import hub_batches_graph

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 128 << 20


def both(gpu, A, sources, AT=None):
    """Scores from the allocation-free entry point and from gk_bc_batch_dev."""
    dA = gpu.upload(A)
    dAT = gpu.upload(AT) if AT is not None else None
    out = gpu.alloc(8 * A.m)
    gpu.bc_batch_dev_noalloc(dA, sources, out, AT=dAT)
    got = out.to_host(np.float64, A.m)
    levels = gpu.last_timing()["iterations"]
    gpu.bc_batch_dev(dA, sources, out, AT=dAT)
    want = out.to_host(np.float64, A.m)
    assert gpu.last_timing()["iterations"] == levels
    out.free()
    dA.free()
    if dAT is not None:
        dAT.free()
    return got, want


def close(got, want):
    err = np.abs(got - want) / np.maximum(np.abs(want), 1.0)
    return float(err.max()) if err.size else 0.0


def test_noalloc_matches_bc_batch_dev_two_calls(gpu):
    G = HostGraph.rmat(14, 16, seed=2)  # 16,384 vertices: 34 MB of scratch plus the levels
    rng = np.random.default_rng(3)
    src = np.sort(rng.choice(np.nonzero(np.diff(G.A.Rst) > 0)[0], size=300, replace=False)).astype(np.int32)
    got, want = both(gpu, G.A, src)
    assert np.array_equal(got, want)
    assert close(got, O.bc_batch(G.A, src)) <= 1e-9
    # second call: a smaller graph and one partial batch, over the first
    # call's leftovers in the scratch array
    H = HostGraph.rmat(10, 8, seed=5)
    src2 = np.arange(0, 40, dtype=np.int32)
    got, want = both(gpu, H.A, src2)
    assert np.array_equal(got, want)
    assert close(got, O.bc_batch(H.A, src2)) <= 1e-9


def test_noalloc_hubs_reached_earlier_in_a_later_batch(gpu):
    # heavy vertices that every source of the second batch reaches before
    # level 2: their frontier words of that level are rewritten, not kept
    # from the first batch (the scratch array is never cleared as a whole)
    G, src = hub_batches_graph()
    got, want = both(gpu, G.A, src)
    assert np.array_equal(got, want)
    assert close(got, O.bc_batch(G.A, src)) <= 1e-9


def test_noalloc_directed_with_transpose(gpu):
    rng = np.random.default_rng(12)
    n = 500
    rst, cols = [0], []
    for i in range(n):
        c = np.sort(rng.choice(np.setdiff1d(np.arange(n), [i]), size=rng.integers(0, 7), replace=False))
        cols.append(c)
        rst.append(rst[-1] + c.size)
    A = SpMtx(n, n, np.array(rst, np.int32), np.concatenate(cols).astype(np.int32))
    rows = np.repeat(np.arange(n), np.diff(A.Rst))
    order = np.lexsort((rows, A.Col))
    trst = np.concatenate([[0], np.cumsum(np.bincount(A.Col, minlength=n))]).astype(np.int32)
    AT = SpMtx(n, n, trst, rows[order].astype(np.int32))
    src = np.arange(0, n, 3, dtype=np.int32)
    got, want = both(gpu, A, src, AT=AT)
    assert np.array_equal(got, want)
    assert close(got, O.brandes_bc(A, src)) <= 1e-9


def test_noalloc_rejects_what_it_cannot_hold(gpu):
    # too many vertices for the scratch array: rejected before any kernel
    n = SCRATCH_BYTES // 2048 + 1
    G = HostGraph.from_edges(n, [1, 2], [0, 1])
    dA = gpu.upload(G.A)
    out = gpu.alloc(8 * n)
    with pytest.raises(GkError, match="scratch array"):
        gpu.bc_batch_dev_noalloc(dA, np.array([0], np.int32), out)
    out.free()
    dA.free()
    # a directed pattern without its transpose
    A = SpMtx(3, 3, np.array([0, 1, 2, 2], np.int32), np.array([1, 2], np.int32))
    dA = gpu.upload(A)
    out = gpu.alloc(8 * 3)
    with pytest.raises(GkError, match="AT or a symmetric A"):
        gpu.bc_batch_dev_noalloc(dA, np.array([0], np.int32), out)
    # the regular entry point still works after the rejections
    gpu.bc_batch_dev(dA, np.array([0], np.int32), out)
    assert np.allclose(out.to_host(np.float64, 3), [0.0, 1.0, 0.0])
    out.free()
    dA.free()


def test_noalloc_symbol_exported():
    assert hasattr(lib(), "gk_bc_batch_dev_noalloc")
