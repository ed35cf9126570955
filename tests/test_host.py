"""Host-side checks that need no GPU: the C-ABI library loads and exports
every entry point include/graphhip.h declares, the generators and the
cleaning pipeline produce valid CSR, the Matrix Market reader follows the
reference's semantics, and the numpy structural helpers match scipy."""
import re

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import ROOT
from graphkokkos_amd import (GraphHipError, HostGraph, SpMtx, build_adjacency_matrix, build_symm_from_lower,
                             extract_lower_triangle, transpose)
from graphkokkos_amd import _lib


def declared_symbols():
    text = (ROOT / "include" / "graphhip.h").read_text()
    return sorted(set(re.findall(r"\b(gk_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_symbols()
    assert len(names) >= 30
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(_lib.exported_symbols())
    assert L.gk_abi_version() == 1


def test_no_gpu_needed_for_errors():
    # argument checks fail before any device work
    with pytest.raises(GraphHipError):
        HostGraph.rmat(0)


def csr_invariants(G: HostGraph):
    A = G.A
    assert A.Rst[0] == 0 and np.all(np.diff(A.Rst) >= 0)
    rows = np.repeat(np.arange(A.m), np.diff(A.Rst))
    assert not np.any(rows == A.Col)                      # no self loops
    for i in range(0, A.m, max(1, A.m // 200)):
        r = A.row(i)
        assert np.all(np.diff(r) > 0)                      # sorted, no duplicates
    S = A.to_scipy()
    assert (S != S.T).nnz == 0                             # symmetric
    L = G.L
    assert np.array_equal(L.to_scipy().toarray() if A.m < 3000 else L.Rst, sp.tril(S, -1).tocsr().toarray()
                          if A.m < 3000 else sp.tril(S, -1).tocsr().indptr)


@pytest.mark.parametrize("make", [
    lambda: HostGraph.rmat(10, 16, seed=1),
    lambda: HostGraph.er(2000, 10.0, seed=2),
    lambda: HostGraph.chung_lu(3000, 2.1, 16.0, seed=3),
    lambda: HostGraph.grid(20),
])
def test_generated_graphs_are_clean(make):
    G = make()
    csr_invariants(G)
    deg = np.diff(G.A.Rst)
    assert np.all(np.diff(deg) <= 0)                       # relabelled by degree descending


def test_generators_are_deterministic():
    a, b = HostGraph.rmat(12, 8, seed=5), HostGraph.rmat(12, 8, seed=5)
    assert np.array_equal(a.A.Col, b.A.Col)
    c = HostGraph.rmat(12, 8, seed=6)
    assert not np.array_equal(a.A.Col, c.A.Col)


def test_grid_sizes():
    for r in (1, 2, 3, 17):
        G = HostGraph.grid(r)
        assert G.edges == 3 * r * r - 4 * r + 1


def test_relabel_is_a_permutation():
    G = HostGraph.rmat(10, 16, seed=3)
    assert np.array_equal(np.sort(G.perm), np.arange(G.n))
    H = HostGraph.rmat(10, 16, seed=3, relabel=False)
    # same graph up to the permutation
    S = H.A.to_scipy().tocoo()
    P = sp.csr_matrix((np.ones(S.nnz), (G.perm[S.row], G.perm[S.col])), shape=S.shape)
    assert (P != G.A.to_scipy()).nnz == 0


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_mtx_reader_reference_semantics(tmp_path):
    # symmetric pattern: mirrored, self loop dropped, duplicate kept,
    # column order = file order
    p = write(tmp_path, "a.mtx", "%%MatrixMarket matrix coordinate pattern symmetric\n% comment\n"
              "4 4 5\n2 1\n3 3\n4 1\n4 2\n2 1\n")
    A = SpMtx.from_file(p)
    assert A.matcode == "MCPS"
    assert A.m == 4 and A.nz == 8
    assert A.row(0).tolist() == [1, 3, 1]
    assert A.row(1).tolist() == [0, 3, 0]
    assert A.row(3).tolist() == [0, 1]
    # general integer with values, 1-based
    p = write(tmp_path, "b.mtx", "%%MatrixMarket matrix coordinate integer general\n3 3 3\n1 3 7\n3 1 -2\n2 2 9\n")
    B = SpMtx.from_file(p)
    assert B.nz == 2 and B.Val.tolist() == [7, -2] and B.matcode == "MCIG"
    # real values
    p = write(tmp_path, "c.mtx", "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 2 0.5\n")
    assert SpMtx.from_file(p).Val.tolist() == [0.5]
    # adjacency + symmetrise-from-lower as the CLI does (src/main.cpp:214-215, 339)
    Adj = build_adjacency_matrix(A)
    assert set(Adj.Val.tolist()) == {1}


@pytest.mark.parametrize("text", [
    "%%MatrixMarket matrix coordinate complex general\n2 2 1\n1 2 1 1\n",
    "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n",
    "%%MatrixMarket matrix coordinate pattern general\n2 3 1\n1 2\n",
    "not a banner\n",
])
def test_mtx_reader_rejects(tmp_path, text):
    with pytest.raises(GraphHipError):
        SpMtx.from_file(write(tmp_path, "x.mtx", text))


def test_structural_helpers_match_scipy():
    rng = np.random.default_rng(0)
    S = sp.random(60, 60, density=0.1, random_state=1, format="csr")
    S = ((S + S.T) > 0).astype(np.int32).tocsr()
    S.setdiag(0)
    S.eliminate_zeros()
    S.sort_indices()
    A = SpMtx(60, 60, S.indptr, S.indices, S.data)
    L = extract_lower_triangle(A)
    Ls = sp.tril(S, -1).tocsr()
    assert np.array_equal(L.Rst, Ls.indptr) and np.array_equal(L.Col, Ls.indices)
    T = transpose(L)
    Ts = Ls.T.tocsr()
    Ts.sort_indices()
    assert np.array_equal(T.Rst, Ts.indptr) and np.array_equal(T.Col, Ts.indices)
    B = build_symm_from_lower(L)
    assert np.array_equal(B.Rst, S.indptr) and np.array_equal(B.Col, S.indices)
    del rng


def test_structural_helpers_on_unsorted_rows():
    # extract_lower_triangle keeps each row's leading run of columns below
    # the diagonal (include/matrix.h:698-726), not every such column
    A = SpMtx(4, 4, np.array([0, 0, 3, 7, 9], np.int32),
              np.array([0, 3, 0, 1, 0, 3, 1, 3, 0], np.int32), np.arange(1, 10, dtype=np.int32))
    L = extract_lower_triangle(A)
    assert L.Rst.tolist() == [0, 0, 1, 3, 3]
    assert L.Col.tolist() == [0, 1, 0] and L.Val.tolist() == [1, 4, 5]
    # build_symm_from_lower merges each row of L with the same row of L^T
    # (include/matrix.h:662-696); an unsorted L row stays in its order
    L = SpMtx(4, 4, np.array([0, 0, 1, 3, 5], np.int32), np.array([0, 1, 0, 2, 0], np.int32),
              np.ones(5, np.int32))
    S = build_symm_from_lower(L)
    assert S.Rst.tolist() == [0, 3, 5, 8, 10]
    assert S.Col.tolist() == [1, 2, 3, 0, 2, 1, 0, 3, 2, 0]


def test_dropin_headers_compile():
    """The C++ drop-in headers (include/graphkokkos) compile warning-free as
    a client program would include them."""
    import subprocess
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror", "-Wno-sign-compare",
                        f"-I{ROOT / 'include'}", str(ROOT / "tests" / "cpp" / "dropin_check.cpp")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_batch_union_is_block_diagonal_copy_of_each_graph():
    from graphkokkos_amd.graph import batch_config
    cfg = batch_config(12, seed=3, n_min=50, n_max=400)
    U = HostGraph.batch(**cfg)
    voff = U.voff
    assert voff[0] == 0 and voff[-1] == U.n and np.all(np.diff(voff) == cfg["n"])
    for g in range(12):
        if cfg["kind"][g] == 0:
            G = HostGraph.er(int(cfg["n"][g]), float(cfg["mean_degree"][g]), seed=int(cfg["seed"][g]))
        else:
            G = HostGraph.chung_lu(int(cfg["n"][g]), float(cfg["exponent"][g]), float(cfg["mean_degree"][g]),
                                   seed=int(cfg["seed"][g]))
        a, b = int(voff[g]), int(voff[g + 1])
        rst = U.A.Rst[a:b + 1] - U.A.Rst[a]
        assert np.array_equal(rst, G.A.Rst)
        assert np.array_equal(U.A.Col[U.A.Rst[a]:U.A.Rst[b]] - a, G.A.Col)
        assert np.array_equal(U.L.Col[U.L.Rst[a]:U.L.Rst[b]] - a, G.L.Col)


def test_batch_config_follows_the_specification():
    from graphkokkos_amd.graph import batch_config
    cfg = batch_config(4096, seed=1)
    n, kind = cfg["n"], cfg["kind"]
    assert n.min() >= 1024 and n.max() <= 65536
    assert abs(np.median(np.log(n)) - 0.5 * (np.log(1024) + np.log(65536))) < 0.1   # log-uniform
    assert (kind == 0).sum() == (kind == 1).sum() == 2048
    d = cfg["mean_degree"][kind == 0]
    assert d.min() >= 8 and d.max() <= 32
    x = cfg["exponent"][kind == 1]
    assert x.min() >= 2.1 and x.max() <= 2.8
