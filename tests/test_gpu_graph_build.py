"""On-device graph construction against the host builder: the same seeded
generators and cleaning must give the same CSR, entry for entry, and the
same relabelling; a device-built graph then feeds the path directly."""
import sys

import numpy as np
import pytest

from graphkokkos_amd import HostGraph
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same(dev, host):
    h = dev.to_host()
    assert (h.m, h.n, h.nz) == (host.m, host.n, host.nz)
    assert np.array_equal(np.asarray(h.Rst), np.asarray(host.Rst))
    assert np.array_equal(np.asarray(h.Col), np.asarray(host.Col))


CASES = [
    ("rmat", dict(scale=1, edge_factor=4, seed=3)),       # mostly self loops and repeats
    ("rmat", dict(scale=8, edge_factor=8, seed=1)),
    ("rmat", dict(scale=12, edge_factor=16, seed=7)),
    ("rmat", dict(scale=16, edge_factor=16, seed=1)),
    ("rmat", dict(scale=14, edge_factor=16, seed=2, a=0.45, b=0.15, c=0.15)),
    ("er", dict(n=2, mean_degree=3.0, seed=1)),
    ("er", dict(n=5000, mean_degree=24.0, seed=3)),
    ("chung_lu", dict(n=20000, exponent=2.0, mean_degree=16.0, seed=5)),
    ("chung_lu", dict(n=3000, exponent=2.7, mean_degree=8.0, seed=9)),
    ("grid", dict(rows=1, cols=1)),                       # no edges
    ("grid", dict(rows=64)),
    ("grid", dict(rows=37, cols=5, diagonal=False)),
]


@pytest.mark.parametrize("relabel", [True, False])
@pytest.mark.parametrize("kind,kw", CASES)
def test_device_build_matches_host(gpu, kind, kw, relabel):
    say("graph build", kind, kw, relabel)
    host = getattr(HostGraph, kind)(**kw, relabel=relabel)
    A, L, perm = getattr(gpu, "dgraph_" + kind)(**kw, relabel=relabel, perm=True)
    same(A, host.A)
    same(L, host.L)
    assert np.array_equal(perm, host.perm)
    A.free()
    L.free()


def test_device_built_graph_feeds_the_path(gpu):
    host = HostGraph.rmat(14, 16, seed=1)
    A, L = gpu.dgraph_rmat(14, 16, seed=1)
    want = O.tc_sandia(host.L)
    assert gpu.tc_count(L) == want
    tr, kmax = gpu.truss_decompose(host.A)
    buf = gpu.alloc(4 * A.nz)
    import ctypes as C
    from graphkokkos_amd import _lib
    k = C.c_int32(0)
    _lib.check(_lib.lib().gk_truss_decompose_dev(gpu.handle, A.handle, buf.ptr, C.byref(k), None),
               "gk_truss_decompose_dev")
    assert int(k.value) == kmax
    assert np.array_equal(buf.to_host(np.uint32, A.nz), tr)


def test_device_build_rmat22_matches_host(gpu):
    """The full-size triangle-counting input (R-MAT scale 22, ef 16, seed 1)."""
    host = HostGraph.rmat(22, 16, seed=1)
    A, L = gpu.dgraph_rmat(22, 16, seed=1)
    same(L, host.L)
    same(A, host.A)
