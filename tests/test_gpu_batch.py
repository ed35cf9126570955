"""Batching: triangle counts and k-trusses of many uneven graphs in one
captured launch sequence (gk_batch_tc_ktruss_dev), against the oracle on
every graph, and hipGraph replays against the eager run."""
import sys

import numpy as np
import pytest

from graphkokkos_amd import GraphHipError, HostGraph, SpMtx
from graphkokkos_amd.graph import batch_config
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def block(M: SpMtx, a: int, b: int) -> SpMtx:
    rst = M.Rst[a:b + 1] - M.Rst[a]
    return SpMtx(b - a, b - a, rst.astype(np.int32), (M.Col[M.Rst[a]:M.Rst[b]] - a).astype(np.int32))


@pytest.fixture(scope="module")
def small_batch():
    cfg = batch_config(48, seed=11, n_min=64, n_max=3000)
    return HostGraph.batch(**cfg)


@pytest.mark.parametrize("k", [3, 4, 5])
def test_batch_matches_oracle_per_graph(gpu, small_batch, k):
    U = small_batch
    say("batch parity", len(U.voff) - 1, "graphs, k =", k)
    d = gpu.upload(U.A)
    r = gpu.batch_tc_ktruss(d, U.voff, k=k, use_graph=True, reps=1, keep=True)
    keep_want, _ = O.ktruss(U.A, k)           # the union's k-truss is the union of the graphs' k-trusses
    assert np.array_equal(r["keep"], keep_want)
    for g in range(len(U.voff) - 1):
        a, b = int(U.voff[g]), int(U.voff[g + 1])
        assert int(r["triangles"][g]) == O.tc_sandia(block(U.L, a, b)), g
        kept = int(keep_want[U.A.Rst[a]:U.A.Rst[b]].sum()) // 2
        assert int(r["truss_edges"][g]) == kept, g
    d.free()


def test_batch_graph_replays_repeat_the_eager_run(gpu, small_batch):
    U = small_batch
    d = gpu.upload(U.A)
    eager = gpu.batch_tc_ktruss(d, U.voff, k=4, use_graph=False, reps=2, digests=True)
    assert eager["digests"][0] == eager["digests"][1]
    # eager work on the same context between two captured runs
    gpu.tc_count(d)
    graph = gpu.batch_tc_ktruss(d, U.voff, k=4, use_graph=True, reps=5, digests=True)
    assert np.all(graph["digests"] == eager["digests"][0]), graph["digests"]
    assert np.array_equal(graph["triangles"], eager["triangles"])
    assert np.array_equal(graph["truss_edges"], eager["truss_edges"])
    gpu.tc_count(d)
    again = gpu.batch_tc_ktruss(d, U.voff, k=4, use_graph=True, reps=3, digests=True)
    assert np.all(again["digests"] == eager["digests"][0])
    d.free()


def test_batch_options_do_not_change_results(gpu, small_batch):
    U = small_batch
    d = gpu.upload(U.A)
    ref = gpu.batch_tc_ktruss(d, U.voff, k=4)
    for o in ({"batch_blocks": 1, "batch_order": 1}, {"batch_blocks": 16}, {"mca_max_mask": 16, "small_max_mask": 32}):
        r = gpu.batch_tc_ktruss(d, U.voff, k=4, opts=o)
        assert np.array_equal(r["triangles"], ref["triangles"]) and np.array_equal(r["truss_edges"], ref["truss_edges"])
    d.free()


def test_batch_rejects_edges_between_graphs(gpu):
    G = HostGraph.grid(6)
    d = gpu.upload(G.A)
    with pytest.raises(GraphHipError, match="block-diagonal"):
        gpu.batch_tc_ktruss(d, np.array([0, 10, G.n]), k=4)
    d.free()
