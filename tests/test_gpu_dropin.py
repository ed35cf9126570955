"""The C++ drop-in headers (include/graphkokkos/*.h) against the oracle.

tests/cpp/dropin_check is a client of libgraphhip.so written the way
GraphKokkos's src/main.cpp uses its own headers: it loads a Matrix Market
file into spMtx<int>, builds the adjacency and symmetric patterns, and runs
mxmm_msa / mxmm_mca through mxmOp<int> pointers, the complement-mask
product mxmm_msa_cmask, the unwired drivers triangle_counting_vertex and
triangle_counting_masked_lu,
triangle_counting_masked_sandia and k_truss (both with the library's
product, which takes the fused GPU entry points, and with a wrapped product,
which takes the reference's loop), the truss decomposition, the supports and
betweenness_centrality_batch.  Integer results must match the oracle
bit-exactly; BC comes back as float, so it is compared within float
rounding (2e-7 relative) of the FP64 oracle.
"""
import json
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT, load_golden
from graphkokkos_amd import SpMtx
from graphkokkos_amd.graph import build_adjacency_matrix, build_symm_from_lower, extract_lower_triangle
from oracle import oracle as O

pytestmark = pytest.mark.gpu

BIN = ROOT / "tests" / "cpp" / "_build" / "dropin_check"


def write_symmetric_lower(path, A: SpMtx) -> None:
    """Lower triangle of a symmetric pattern as a 'pattern symmetric' file."""
    rows = np.repeat(np.arange(A.m), np.diff(A.Rst))
    keep = A.Col < rows
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate pattern symmetric\n")
        f.write(f"{A.m} {A.n} {int(keep.sum())}\n")
        for i, j in zip(rows[keep], A.Col[keep]):
            f.write(f"{i + 1} {j + 1}\n")


def write_scrambled_general(path, A: SpMtx, seed: int) -> None:
    """Both directions of every edge in random order, some entries repeated,
    some self loops, integer values: rows come out unsorted with duplicate
    columns, which the masked product must accept as the reference's MSA
    does."""
    rng = np.random.default_rng(seed)
    rows = np.repeat(np.arange(A.m), np.diff(A.Rst))
    e = np.stack([rows, A.Col], 1)
    dup = e[rng.choice(len(e), len(e) // 10, replace=False)]
    loops = np.repeat(rng.choice(A.m, 5, replace=False)[:, None], 2, 1)
    e = np.concatenate([e, dup, loops])
    e = e[rng.permutation(len(e))]
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate integer general\n")
        f.write(f"{A.m} {A.n} {len(e)}\n")
        for i, j in e:
            f.write(f"{i + 1} {j + 1} {int(rng.integers(1, 9))}\n")


def run_check(mtx, tmp_path, k, batch, loaded_only=False):
    assert BIN.exists(), f"{BIN} is missing: build with `python -m graphkokkos_amd.build`"
    out = tmp_path / "out.json"
    args = [str(BIN), str(mtx), str(out), str(k), str(batch)] + (["loaded"] if loaded_only else [])
    print(*args[1:], flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    sys.stdout.write(r.stdout[-2000:])
    assert r.returncode == 0, r.stderr[-4000:]
    return json.loads(out.read_text())


def check_against_oracle(res, mtx, k, batch, loaded_only=False):
    gr = SpMtx.from_file(str(mtx))
    A = build_adjacency_matrix(gr)
    assert np.array_equal(res["loaded_rst"], A.Rst) and np.array_equal(res["loaded_col"], A.Col)
    Ap = SpMtx(A.m, A.n, A.Rst, A.Col, np.ones(A.nz, np.int32))
    assert np.array_equal(res["mxm_loaded"], O.dense_masked_product(Ap, Ap, Ap))
    rst, col, val = O.dense_cmask_product(Ap, Ap, Ap)
    assert np.array_equal(res["cmask_rst"], rst) and np.array_equal(res["cmask_col"], col)
    assert np.array_equal(res["cmask_val"], val)
    want_bc = O.bc_batch(Ap, np.arange(min(A.m, batch)))
    got_bc = np.asarray(res["bc"], np.float64)
    assert np.all(np.abs(got_bc - want_bc) <= 2e-7 * np.maximum(np.abs(want_bc), 1.0))
    assert res["bc_nonsquare_throws"] == 1
    if loaded_only:
        return
    S = build_symm_from_lower(extract_lower_triangle(A))
    S = SpMtx(S.m, S.n, S.Rst, S.Col, np.ones(S.nz, np.int32))
    assert np.array_equal(res["sym_rst"], S.Rst) and np.array_equal(res["sym_col"], S.Col)
    assert np.array_equal(res["mxm_sym"], O.mxmm(S, S, S))
    assert res["mxm_return_form_equal"] == 1
    L = extract_lower_triangle(S)
    t = O.tc_sandia(L)
    assert res["tc_fused"] == t and res["tc_loop"] == t
    assert res["tc_lu"] == t and res["tc_lu_loop"] == t
    sup = O.support(S)
    vrows = np.repeat(np.arange(S.m), np.diff(S.Rst))
    assert np.array_equal(res["tc_vertex"], np.bincount(vrows, weights=sup, minlength=S.m).astype(np.int64) // 2)
    assert res["tc_vertex_loop_equal"] == 1
    keep, _ = O.ktruss(S, k)
    rows = np.repeat(np.arange(S.m), np.diff(S.Rst))
    want_rst = np.concatenate([[0], np.cumsum(np.bincount(rows[keep], minlength=S.m))])
    assert np.array_equal(res["ktruss_rst"], want_rst)
    assert np.array_equal(res["ktruss_col"], S.Col[keep])
    assert res["ktruss_loop_equal"] == 1
    truss, kmax = O.truss_peel(S)
    assert np.array_equal(res["truss"], truss) and res["kmax"] == kmax
    assert np.array_equal(res["support"], O.support(S))


@pytest.mark.parametrize("name,k", [("karate", 4), ("plc80", 5), ("grid5", 3)])
def test_dropin_symmetric_file(tmp_path, name, k):
    A, _ = load_golden(name)
    mtx = tmp_path / f"{name}.mtx"
    write_symmetric_lower(mtx, A)
    res = run_check(mtx, tmp_path, k, 16)
    check_against_oracle(res, mtx, k, 16)


@pytest.mark.parametrize("name,seed", [("er60", 1), ("karate", 2)])
def test_dropin_scrambled_general_file(tmp_path, name, seed):
    A, _ = load_golden(name)
    mtx = tmp_path / f"{name}_scrambled.mtx"
    write_scrambled_general(mtx, A, seed)
    res = run_check(mtx, tmp_path, 3, 8, loaded_only=True)
    check_against_oracle(res, mtx, 3, 8, loaded_only=True)
