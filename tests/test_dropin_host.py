"""Host-only parts of the C++ drop-in header graphkokkos/matrix.h that never
reach the GPU: denseMtx<ValT> (GraphKokkos include/matrix.h:548-630), the
free deep_copy (:536-546), convertType (:632-641) and
extract_upper_triangle (:728-755), driven by tests/cpp/dense_check.cpp and
compared with numpy's dense and triangular views of the same CSR."""
import json
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT

SRC = ROOT / "tests" / "cpp" / "dense_check.cpp"


@pytest.fixture(scope="module")
def dense_check(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("dense") / "dense_check"
    subprocess.run([gxx, "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-Wno-sign-compare",
                    f"-I{ROOT / 'include'}", str(SRC), "-o", str(out)], check=True)
    return out


def random_csr(rng, m, n, density, sorted_rows=True):
    rows, cols = [], []
    for i in range(m):
        c = np.nonzero(rng.random(n) < density)[0]
        if not sorted_rows:
            rng.shuffle(c)
        rows.append(np.full(c.size, i))
        cols.append(c)
    col = np.concatenate(cols).astype(np.int64)
    rst = np.concatenate([[0], np.cumsum([c.size for c in cols])]).astype(np.int64)
    val = rng.integers(-50, 50, col.size)
    return rst, col, val


def run(binary, m, n, rst, col, val):
    text = f"{m} {n} {col.size}\n" + " ".join(map(str, rst)) + "\n" + " ".join(map(str, col)) + "\n" + \
        " ".join(map(str, val)) + "\n"
    r = subprocess.run([str(binary)], input=text, capture_output=True, text=True, check=True)
    out = {}
    for line in r.stdout.splitlines():
        out.update(json.loads(line))
    return out


def dense(m, n, rst, col, val):
    D = np.zeros((m, n), np.int64)
    for i in range(m):
        D[i, col[rst[i]:rst[i + 1]]] = val[rst[i]:rst[i + 1]]
    return D


@pytest.mark.parametrize("m,n,density,sorted_rows", [(1, 1, 1.0, True), (7, 5, 0.4, True), (12, 12, 0.3, False),
                                                     (30, 40, 0.1, True), (5, 9, 0.0, True)])
def test_dense_and_triangles(dense_check, m, n, density, sorted_rows):
    rng = np.random.default_rng(m * 100 + n)
    rst, col, val = random_csr(rng, m, n, density, sorted_rows)
    out = run(dense_check, m, n, rst, col, val)
    D = dense(m, n, rst, col, val)
    assert out["dense_from_sparse"] == D.ravel().tolist()
    assert out["assign_from_sparse"] == D.ravel().tolist()
    assert out["equal_after_assign"] == 1
    assert out["zeros"] == [0] * 12
    assert out["copy_shares"] == 1
    # resize to (m + 1) x n keeps the flat array's first m*n values
    assert out["resized"] == D.ravel().tolist() + [0] * n
    assert out["moved"] == [m + 1, n, 0]

    cp = out["deep_copy"]
    assert (cp["m"], cp["n"], cp["nz"]) == (m, n, col.size)
    assert cp["rst"] == rst.tolist() and cp["col"] == col.tolist()
    if col.size:
        assert out["deep_copy_independent"] == 1 and cp["val"] == [val[0] + 1000] + val[1:].tolist()
    if col.size != 1:
        assert out["alias_kept"] == 1

    cv = out["convert"]
    assert cv["rst"] == rst.tolist() and cv["col"] == col.tolist() and cv["val"] == [float(v) for v in val]

    # leading run of columns below (lower) / on-or-below (upper: the rest) the diagonal
    for name, below in (("lower", lambda c, i: c < i), ("upper", lambda c, i: c <= i)):
        t = out[name]
        want_rst, want_col, want_val = [0], [], []
        for i in range(m):
            r = col[rst[i]:rst[i + 1]]
            v = val[rst[i]:rst[i + 1]]
            p = 0
            while p < r.size and below(r[p], i):
                p += 1
            sel = slice(0, p) if name == "lower" else slice(p, r.size)
            want_col += r[sel].tolist()
            want_val += v[sel].tolist()
            want_rst.append(len(want_col))
        assert t["rst"] == want_rst and t["col"] == want_col and t["val"] == want_val, name
        if sorted_rows:
            # on sorted rows these are the strict triangles of the dense matrix
            T = dense(m, n, np.asarray(t["rst"]), np.asarray(t["col"], np.int64), np.asarray(t["val"]))
            assert np.array_equal(T, np.tril(D, -1) if name == "lower" else np.triu(D, 1))
