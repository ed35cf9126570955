"""ORACLE -- test infrastructure only.

Python access to the CPU restatement of GraphKokkos's algorithms
(oracle/oracle.cpp, built into oracle/liboracle.so by oracle/Makefile) plus
small pure-Python brute-force checkers used to pin it.  Only tests/, the
smoke check in __graft_entry__.py and the CPU-baseline leg of bench.py use
this module; the product library never does.

Restated reference routines (GraphKokkos tree):
  mxmm(..)             include/matrix_la.h:493-560 (MCA), :576-635 / :696-731 (MSA)
  tc_sandia(..)        src/main.cpp:69-90
  support(..)          src/main.cpp:228-233 ("mxm" action, S<A>=A*A)
  ktruss(..)           src/main.cpp:93-141
  truss_decompose(..)  src/main.cpp:93-141 looped over k
  truss_peel(..)       independent peeling algorithm (cross-check)
  bc_batch(..)         include/betweenness_centrality.h:182-242, 313-506

Brute-force checkers (independent of the restatement):
  dense_masked_product  C = (A @ B) .* pattern(M) with numpy integers
  brandes_bc            textbook Brandes by BFS from every source, FP64
  brute_truss           truss numbers by the definition (repeat: drop edges
                        with fewer than k-2 triangles inside the subgraph)
"""
from __future__ import annotations

import ctypes as C
import subprocess
from collections import deque
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

_i32 = C.POINTER(C.c_int32)
_u32 = C.POINTER(C.c_uint32)
_u8 = C.POINTER(C.c_uint8)
_f64 = C.POINTER(C.c_double)
_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        L.or_num_threads.restype = C.c_int
        L.or_set_threads.argtypes = [C.c_int]
        L.or_mxmm.restype = C.c_int
        L.or_mxmm.argtypes = [C.c_int, C.c_int] + [C.c_int64, C.c_int64, C.c_int64, _i32, _i32, _i32] * 2 + \
            [C.c_int64, C.c_int64, C.c_int64, _i32, _i32, _i32]
        L.or_mxmm_cmask.restype = C.c_int64
        L.or_mxmm_cmask.argtypes = [C.c_int] + [C.c_int64, C.c_int64, C.c_int64, _i32, _i32, _i32] * 2 + \
            [C.c_int64, C.c_int64, C.c_int64, _i32, _i32, _i32, _i32, _i32]
        L.or_tc_sandia.restype = C.c_int64
        L.or_tc_sandia.argtypes = [C.c_int, C.c_int, C.c_int64, C.c_int64, _i32, _i32]
        L.or_support.restype = C.c_int
        L.or_support.argtypes = [C.c_int, C.c_int, C.c_int64, C.c_int64, _i32, _i32, _i32]
        L.or_ktruss.restype = C.c_int64
        L.or_ktruss.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int64, C.c_int64, _i32, _i32, _u8]
        L.or_truss_decompose.restype = C.c_int
        L.or_truss_decompose.argtypes = [C.c_int, C.c_int, C.c_int64, C.c_int64, _i32, _i32, _u32,
                                         C.POINTER(C.c_int64)]
        L.or_truss_peel.restype = C.c_int
        L.or_truss_peel.argtypes = [C.c_int64, C.c_int64, _i32, _i32, _u32]
        L.or_bc_batch.restype = C.c_int
        L.or_bc_batch.argtypes = [C.c_int, C.c_int64, C.c_int64, _i32, _i32, _i32, C.c_int64, C.c_int64, _f64]
        _lib = L
    return _lib


def _p(a, t=_i32):
    return None if a is None else np.ascontiguousarray(a).ctypes.data_as(t)


def _csr(a):
    """(m, n, nz, Rst, Col, Val-or-None) from an object with Rst/Col/Val."""
    rst = np.ascontiguousarray(a.Rst, np.int32)
    col = np.ascontiguousarray(a.Col, np.int32)
    val = None if getattr(a, "Val", None) is None else np.ascontiguousarray(a.Val, np.int32)
    return a.m, a.n, int(rst[-1]), rst, col, val


def threads() -> int:
    return lib().or_num_threads()


def set_threads(n: int) -> None:
    lib().or_set_threads(n)


MSA, MCA, MSA_STALE = 0, 1, 2


def mxmm(A, B, M, accum: int = MSA, parallel: bool = True) -> np.ndarray:
    """Values of C<M> = A*B at M's pattern (int32, wrap-around)."""
    am, an, anz, ar, ac, av = _csr(A)
    bm, bn, bnz, br, bc, bv = _csr(B)
    mm, mn, mnz, mr, mc, _ = _csr(M)
    out = np.zeros(mnz, np.int32)
    rc = lib().or_mxmm(accum, int(parallel), am, an, anz, _p(ar), _p(ac), _p(av), bm, bn, bnz, _p(br), _p(bc),
                       _p(bv), mm, mn, mnz, _p(mr), _p(mc), _p(out))
    if rc != 0:
        raise ValueError("shape mismatch")
    return out


def mxmm_cmask(A, B, M, parallel: bool = True) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """C<not M> = A*B (mxmm_msa_cmask): (rst, col, val) of C, rows in column
    order, explicit zeros kept, int32 wrap-around."""
    am, an, anz, ar, ac, av = _csr(A)
    bm, bn, bnz, br, bc, bv = _csr(B)
    mm, mn, mnz, mr, mc, _ = _csr(M)
    rst = np.zeros(am + 1, np.int32)
    nz = lib().or_mxmm_cmask(int(parallel), am, an, anz, _p(ar), _p(ac), _p(av), bm, bn, bnz, _p(br), _p(bc), _p(bv),
                             mm, mn, mnz, _p(mr), _p(mc), _p(rst), None, None)
    if nz < 0:
        raise ValueError("shape mismatch")
    col = np.zeros(max(nz, 1), np.int32)
    val = np.zeros(max(nz, 1), np.int32)
    lib().or_mxmm_cmask(int(parallel), am, an, anz, _p(ar), _p(ac), _p(av), bm, bn, bnz, _p(br), _p(bc), _p(bv),
                        mm, mn, mnz, _p(mr), _p(mc), _p(rst), _p(col), _p(val))
    return rst, col[:nz], val[:nz]


def tc_sandia(L, accum: int = MSA, parallel: bool = True) -> int:
    m, _, nz, r, c, _ = _csr(L)
    return int(lib().or_tc_sandia(accum, int(parallel), m, nz, _p(r), _p(c)))


def support(A, accum: int = MSA, parallel: bool = True) -> np.ndarray:
    m, _, nz, r, c, _ = _csr(A)
    out = np.zeros(nz, np.int32)
    lib().or_support(accum, int(parallel), m, nz, _p(r), _p(c), _p(out))
    return out


def ktruss(A, k: int, accum: int = MSA, parallel: bool = True) -> tuple[np.ndarray, int]:
    """keep-mask over A's entries of the k-truss, and the reference's
    iteration count."""
    m, _, nz, r, c, _ = _csr(A)
    keep = np.zeros(nz, np.uint8)
    it = lib().or_ktruss(accum, int(parallel), k, m, nz, _p(r), _p(c), _p(keep, _u8))
    return keep.astype(bool), int(it)


def truss_decompose(A, accum: int = MSA, parallel: bool = True) -> tuple[np.ndarray, int, int]:
    m, _, nz, r, c, _ = _csr(A)
    out = np.zeros(nz, np.uint32)
    iters = C.c_int64(0)
    kmax = lib().or_truss_decompose(accum, int(parallel), m, nz, _p(r), _p(c), _p(out, _u32), C.byref(iters))
    return out, int(kmax), int(iters.value)


def truss_peel(A) -> tuple[np.ndarray, int]:
    m, _, nz, r, c, _ = _csr(A)
    out = np.zeros(nz, np.uint32)
    kmax = lib().or_truss_peel(m, nz, _p(r), _p(c), _p(out, _u32))
    return out, int(kmax)


def bc_batch(A, sources, fp64: bool = True, batch: int = 256) -> np.ndarray:
    m, _, nz, r, c, _ = _csr(A)
    src = np.ascontiguousarray(sources, np.int32)
    out = np.zeros(m, np.float64)
    lib().or_bc_batch(int(fp64), m, nz, _p(r), _p(c), _p(src), len(src), batch, _p(out, _f64))
    return out


# ---- brute force (tiny graphs) -------------------------------------------

def dense_masked_product(A, B, M) -> np.ndarray:
    """(A @ B) .* pattern(M) by dense int64 products, reduced mod 2^32 to int32."""
    def dense(X, val=True):
        D = np.zeros((X.m, X.n), np.int64)
        v = np.ones(int(X.Rst[-1]), np.int64) if (getattr(X, "Val", None) is None or not val) else \
            np.asarray(X.Val, np.int64)
        rows = np.repeat(np.arange(X.m), np.diff(X.Rst))
        np.add.at(D, (rows, np.asarray(X.Col)), v)
        return D
    P = dense(A) @ dense(B)
    rows = np.repeat(np.arange(M.m), np.diff(M.Rst))
    return (P[rows, np.asarray(M.Col)] & 0xFFFFFFFF).astype(np.uint32).view(np.int32)


def dense_cmask_product(A, B, M) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """C<not M> = A*B by brute force: the entries some product term reaches
    (a boolean product of the patterns), minus M's pattern, with the int64
    products reduced mod 2^32; rows in column order."""
    def dense(X, val=True):
        D = np.zeros((X.m, X.n), np.int64)
        v = np.ones(int(X.Rst[-1]), np.int64) if (getattr(X, "Val", None) is None or not val) else \
            np.asarray(X.Val, np.int64)
        rows = np.repeat(np.arange(X.m), np.diff(X.Rst))
        np.add.at(D, (rows, np.asarray(X.Col)), v)
        return D
    reach = (dense(A, False) != 0).astype(np.int64) @ (dense(B, False) != 0).astype(np.int64) > 0
    Mp = dense(M, False) != 0
    keep = reach & ~Mp
    P = dense(A) @ dense(B)
    rows, cols = np.nonzero(keep)
    rst = np.concatenate([[0], np.cumsum(keep.sum(axis=1))]).astype(np.int32)
    return rst, cols.astype(np.int32), (P[rows, cols] & 0xFFFFFFFF).astype(np.uint32).view(np.int32)


def _adj(A) -> list[list[int]]:
    return [list(map(int, A.Col[A.Rst[i]:A.Rst[i + 1]])) for i in range(A.m)]


def brandes_bc(A, sources) -> np.ndarray:
    """sum over sources s of delta_s(v) (directed edges i -> Col), FP64,
    textbook Brandes (BFS, sigma counts, reverse accumulation)."""
    adj = _adj(A)
    n = A.m
    bc = np.zeros(n)
    for s in sources:
        s = int(s)
        sigma = [0.0] * n
        dist = [-1] * n
        sigma[s], dist[s] = 1.0, 0
        order, q = [], deque([s])
        while q:
            v = q.popleft()
            order.append(v)
            for w in adj[v]:
                if dist[w] < 0:
                    dist[w] = dist[v] + 1
                    q.append(w)
                if dist[w] == dist[v] + 1:
                    sigma[w] += sigma[v]
        delta = [0.0] * n
        for w in reversed(order):
            for x in adj[w]:
                if dist[x] == dist[w] + 1:
                    delta[w] += sigma[w] / sigma[x] * (1.0 + delta[x])
            if w != s:
                bc[w] += delta[w]
    return bc


def brute_truss(A) -> np.ndarray:
    """Truss number per entry of a symmetric pattern, by the definition."""
    adj = [set(r) for r in _adj(A)]
    edges = {(i, j) for i in range(A.m) for j in adj[i] if j < i}
    alive = set(edges)
    truss = {e: 2 for e in edges}
    k = 3
    while alive:
        changed = True
        while changed:
            changed = False
            nb = {v: set() for v in range(A.m)}
            for (i, j) in alive:
                nb[i].add(j)
                nb[j].add(i)
            drop = [e for e in alive if len(nb[e[0]] & nb[e[1]]) < k - 2]
            if drop:
                changed = True
                for e in drop:
                    alive.discard(e)
                    truss[e] = k - 1
        k += 1
    out = np.zeros(int(A.Rst[-1]), np.uint32)
    for i in range(A.m):
        for p in range(A.Rst[i], A.Rst[i + 1]):
            j = int(A.Col[p])
            out[p] = truss[(max(i, j), min(i, j))]
    return out
