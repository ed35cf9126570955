"""GPU checks at full size, where running the oracle inside the test would
take minutes: triangle counts and the truss-number histogram/checksum of the
seeded R-MAT scale-20 and scale-22 graphs against summaries the oracle wrote
(tests/golden/make_large.py), plus size-independent invariants:
sum of supports = 3 * triangles, truss(e) <= support(e) + 2, every
accumulator gives the same count."""
import json

import numpy as np
import pytest

from conftest import GOLDEN
from graphkokkos_amd import (ACC_DENSE_GLOBAL, ACC_HASH_GLOBAL, ACC_MCA_WAVE, ACC_MSA_LDS, SEL_FORCED,
                             HostGraph)

pytestmark = pytest.mark.gpu


def summary():
    return json.loads((GOLDEN / "rmat_large.json").read_text())


def checksum(a):
    w = (np.arange(a.size, dtype=np.uint64) % np.uint64(1000003)) + np.uint64(1)
    return int(np.sum(a.astype(np.uint64) * w, dtype=np.uint64))


@pytest.fixture(scope="module")
def g20():
    return HostGraph.rmat(20, 16, seed=1)


def test_rmat20_triangles_every_accumulator(gpu, g20):
    want = summary()["rmat20"]
    assert g20.edges == want["edges"]
    d = gpu.upload(g20.L)
    assert gpu.tc_count(d) == want["triangles"]
    for acc in (ACC_MSA_LDS, ACC_MCA_WAVE, ACC_HASH_GLOBAL, ACC_DENSE_GLOBAL):
        assert gpu.tc_count(d, {"selector": SEL_FORCED, "forced_accum": acc}) == want["triangles"], acc


def test_rmat20_truss_decomposition(gpu, g20):
    want = summary()["rmat20"]
    d = gpu.upload(g20.A)
    truss, kmax = gpu.truss_decompose(d)
    assert kmax == want["kmax"]
    assert np.bincount(truss).tolist() == want["truss_hist"]
    assert checksum(truss) == want["truss_checksum"]
    sup = gpu.ktruss_support(g20.A)
    assert int(sup.sum(dtype=np.int64)) == 6 * want["triangles"]
    assert np.all(truss.astype(np.int64) <= sup.astype(np.int64) + 2)


def test_rmat22_triangles(gpu):
    want = summary()["rmat22"]
    G = HostGraph.rmat(22, 16, seed=1)
    assert G.edges == want["edges"]
    assert gpu.tc_count(G.L) == want["triangles"]
