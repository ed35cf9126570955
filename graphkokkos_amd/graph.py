"""Host-side mirror of GraphKokkos's containers and graph operations.

``SpMtx`` mirrors ``spMtx<ValT>`` (include/matrix.h:33-208): CSR arrays
``Rst`` (m+1), ``Col`` (nz), ``Val`` (nz) plus ``m``, ``n``, ``nz`` and the
Matrix Market typecode ``matcode``.  Graph construction (generators,
cleaning, relabelling, the Matrix Market reader) runs in libgraphhip.so on
host threads; the structural helpers below restate the reference's
``extract_lower_triangle`` / ``build_symm_from_lower`` /
``build_adjacency_matrix`` / ``transpose`` with numpy.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import GkCsr, check, lib


def _ptr(a: np.ndarray | None, ct=C.c_int32):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ct))


@dataclass
class SpMtx:
    """CSR matrix with the reference's member names (include/matrix.h:36-42)."""

    m: int = 0
    n: int = 0
    Rst: np.ndarray = field(default_factory=lambda: np.zeros(1, np.int32))
    Col: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    Val: np.ndarray | None = None
    matcode: str = "MCPG"

    def __post_init__(self):
        self.Rst = np.ascontiguousarray(self.Rst, dtype=np.int32)
        self.Col = np.ascontiguousarray(self.Col, dtype=np.int32)
        if self.Val is not None:
            self.Val = np.ascontiguousarray(self.Val)
        if len(self.Rst) != self.m + 1:
            raise ValueError(f"Rst has {len(self.Rst)} entries, expected m+1 = {self.m + 1}")
        if int(self.Rst[-1]) != len(self.Col):
            raise ValueError("Rst[m] must equal len(Col)")

    @property
    def nz(self) -> int:
        return int(self.Rst[-1])

    def row(self, i: int) -> np.ndarray:
        return self.Col[self.Rst[i]:self.Rst[i + 1]]

    def values_i32(self) -> np.ndarray | None:
        """int32 values, or None for a pattern matrix whose values are all 1."""
        if self.Val is None:
            return None
        v = np.ascontiguousarray(self.Val, dtype=np.int32)
        if v.size and np.all(v == 1):
            return None
        return v

    def gk(self, with_values: bool = True) -> GkCsr:
        """A gk_csr view; the arrays stay owned by this object."""
        v = self.values_i32() if with_values else None
        self._keep = v  # keep the int32 copy alive while the struct is used
        return GkCsr(self.m, self.n, self.nz, _ptr(self.Rst), _ptr(self.Col), _ptr(v))

    def to_scipy(self):
        import scipy.sparse as sp
        val = np.ones(self.nz, np.int64) if self.Val is None else self.Val
        return sp.csr_matrix((val, self.Col, self.Rst), shape=(self.m, self.n))

    def __eq__(self, other) -> bool:  # operator== (include/matrix.h:165-175)
        if not isinstance(other, SpMtx):
            return NotImplemented
        va = np.ones(self.nz) if self.Val is None else self.Val
        vb = np.ones(other.nz) if other.Val is None else other.Val
        return (self.m == other.m and self.n == other.n and self.nz == other.nz
                and np.array_equal(self.Rst, other.Rst) and np.array_equal(self.Col, other.Col)
                and np.array_equal(va, vb))

    @classmethod
    def from_file(cls, path: str, fmt: str = "mtx") -> "SpMtx":
        """spMtx(filename, format) (include/matrix.h:44-61): "mtx"
        (read_mtx_to_crs :259-395), "crs", "bin", "graph", "rmat"
        (:396-533; deviations in include/graphhip.h)."""
        h = C.c_void_p()
        check(lib().gk_matrix_read(str(path).encode(), fmt.encode(), C.byref(h)), "gk_matrix_read")
        try:
            csr = GkCsr()
            val = _lib.f64p()
            code = C.create_string_buffer(4)
            check(lib().gk_mtx_view(h, C.byref(csr), C.byref(val), code), "gk_mtx_view")
            m, nz = csr.m, csr.nz
            rst = np.ctypeslib.as_array(csr.rst, (m + 1,)).copy()
            col = np.ctypeslib.as_array(csr.col, (nz,)).copy() if nz else np.zeros(0, np.int32)
            v = np.ctypeslib.as_array(val, (nz,)).copy() if nz else np.zeros(0)
            matcode = code.raw.decode()
            if matcode[2] in "IP":
                v = v.astype(np.int32)
            return cls(int(m), int(csr.n), rst, col, v, matcode)
        finally:
            lib().gk_mtx_free(h)


def write_matrix(a: SpMtx, path: str, fmt: str = "bin", values=None) -> None:
    """write_crs_to_bin / write_crs_to_mtx (include/matrix.h:209-257) and
    the crs text form.  ``values`` (int32 per entry, e.g. supports or truss
    numbers) default to the matrix's own int32 values; a pattern matrix is
    written as a pattern (bin / crs: values 1)."""
    csr = a.gk(with_values=False)
    v = values if values is not None else a.values_i32()
    vp = None
    if v is not None:
        v = np.ascontiguousarray(v, np.int32)
        if len(v) != a.nz:
            raise ValueError(f"{len(v)} values for {a.nz} entries")
        vp = v.ctypes.data_as(C.POINTER(C.c_int32))
    check(lib().gk_matrix_write(str(path).encode(), fmt.encode(), C.byref(csr), vp, None), "gk_matrix_write")


def build_adjacency_matrix(g: SpMtx) -> SpMtx:
    """Pattern with every value 1 (include/matrix.h:643-660)."""
    return SpMtx(g.m, g.n, g.Rst.copy(), g.Col.copy(), np.ones(g.nz, np.int32), g.matcode[:2] + "I" + g.matcode[3])


def _rows_sorted(a: SpMtx) -> bool:
    rows = np.repeat(np.arange(a.m, dtype=np.int64), np.diff(a.Rst))
    same = rows[1:] == rows[:-1]
    return bool(np.all(a.Col[1:][same] > a.Col[:-1][same]))


def extract_lower_triangle(g: SpMtx) -> SpMtx:
    """Each row's leading run of columns below the diagonal; on sorted rows
    that is the strictly lower part (include/matrix.h:698-726)."""
    rst = np.asarray(g.Rst, np.int64)
    rows = np.repeat(np.arange(g.m, dtype=np.int64), np.diff(rst))
    stop = np.where(g.Col >= rows, np.arange(g.nz, dtype=np.int64), np.iinfo(np.int64).max)
    first = rst[1:].copy()
    nonempty = np.flatnonzero(np.diff(rst) > 0)
    if nonempty.size:
        first[nonempty] = np.minimum(first[nonempty], np.minimum.reduceat(stop, rst[nonempty]))
    lens = first - rst[:-1]
    keep_rst = np.zeros(g.m + 1, np.int64)
    np.cumsum(lens, out=keep_rst[1:])
    sel = np.arange(g.nz, dtype=np.int64) < np.repeat(first, np.diff(rst))
    val = None if g.Val is None else g.Val[sel]
    return SpMtx(g.m, g.n, keep_rst.astype(np.int32), g.Col[sel], val, g.matcode)


def transpose(a: SpMtx) -> SpMtx:
    """Counting-sort transpose, rows come out sorted (include/matrix_la.h:155-179)."""
    cnt = np.bincount(a.Col, minlength=a.n).astype(np.int64)
    rst = np.zeros(a.n + 1, np.int64)
    np.cumsum(cnt, out=rst[1:])
    rows = np.repeat(np.arange(a.m, dtype=np.int32), np.diff(a.Rst))
    order = np.lexsort((rows, a.Col))
    val = None if a.Val is None else a.Val[order]
    return SpMtx(a.n, a.m, rst.astype(np.int32), rows[order], val, a.matcode)


def build_symm_from_lower(low: SpMtx) -> SpMtx:
    """Row-wise merge of L and L^T (include/matrix.h:662-696): an entry of L
    is taken while its column is below the next column of L^T's row."""
    up = transpose(low)
    lv = low.Val if low.Val is not None else np.ones(low.nz, np.int32)
    uv = up.Val if up.Val is not None else np.ones(up.nz, np.int32)
    if _rows_sorted(low):
        rows = np.concatenate([np.repeat(np.arange(up.m), np.diff(up.Rst)), np.repeat(np.arange(low.m), np.diff(low.Rst))])
        cols = np.concatenate([up.Col, low.Col])
        vals = np.concatenate([uv, lv])
        order = np.lexsort((cols, rows))   # stable: on equal columns L^T's entry first, as the merge takes
        rst = np.zeros(low.m + 1, np.int64)
        np.cumsum(np.bincount(rows, minlength=low.m), out=rst[1:])
        cols, vals = cols[order], vals[order]
    else:
        rst = np.zeros(low.m + 1, np.int64)
        cols = np.zeros(2 * low.nz, np.int32)
        vals = np.zeros(2 * low.nz, lv.dtype)
        w = 0
        for i in range(low.m):
            a, ae = int(low.Rst[i]), int(low.Rst[i + 1])
            b, be = int(up.Rst[i]), int(up.Rst[i + 1])
            while a < ae or b < be:
                if b >= be or (a < ae and low.Col[a] < up.Col[b]):
                    cols[w], vals[w] = low.Col[a], lv[a]
                    a += 1
                else:
                    cols[w], vals[w] = up.Col[b], uv[b]
                    b += 1
                w += 1
            rst[i + 1] = w
    return SpMtx(low.m, low.n, rst.astype(np.int32), cols.astype(np.int32),
                 None if low.Val is None else vals, low.matcode)


class HostGraph:
    """A cleaned undirected simple graph built by libgraphhip.so: symmetric
    pattern ``A`` (sorted rows) and its strict lower part ``L``."""

    def __init__(self, handle: C.c_void_p, desc: str):
        self._h = handle
        self.desc = desc
        A, L = GkCsr(), GkCsr()
        check(lib().gk_hgraph_view(self._h, C.byref(A), C.byref(L)), "gk_hgraph_view")
        self.n = int(A.m)
        self.A = SpMtx(self.n, self.n, np.ctypeslib.as_array(A.rst, (self.n + 1,)).copy(),
                       np.ctypeslib.as_array(A.col, (A.nz,)).copy() if A.nz else np.zeros(0, np.int32))
        self.L = SpMtx(self.n, self.n, np.ctypeslib.as_array(L.rst, (self.n + 1,)).copy(),
                       np.ctypeslib.as_array(L.col, (L.nz,)).copy() if L.nz else np.zeros(0, np.int32))
        self.perm = np.zeros(self.n, np.int32)
        check(lib().gk_hgraph_perm(self._h, _ptr(self.perm)), "gk_hgraph_perm")
        lib().gk_hgraph_free(self._h)
        self._h = None

    @property
    def edges(self) -> int:
        return self.L.nz

    @classmethod
    def rmat(cls, scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
             seed: int = 1, relabel: bool = True) -> "HostGraph":
        h = C.c_void_p()
        check(lib().gk_hgraph_rmat(scale, edge_factor, a, b, c, seed, int(relabel), C.byref(h)), "gk_hgraph_rmat")
        return cls(h, f"rmat-s{scale}-ef{edge_factor}-seed{seed}")

    @classmethod
    def er(cls, n: int, mean_degree: float, seed: int = 1, relabel: bool = True) -> "HostGraph":
        h = C.c_void_p()
        check(lib().gk_hgraph_er(n, mean_degree, seed, int(relabel), C.byref(h)), "gk_hgraph_er")
        return cls(h, f"er-n{n}-d{mean_degree}-seed{seed}")

    @classmethod
    def chung_lu(cls, n: int, exponent: float, mean_degree: float = 16.0, seed: int = 1,
                 relabel: bool = True) -> "HostGraph":
        h = C.c_void_p()
        check(lib().gk_hgraph_chung_lu(n, exponent, mean_degree, seed, int(relabel), C.byref(h)),
              "gk_hgraph_chung_lu")
        return cls(h, f"chunglu-n{n}-g{exponent}-d{mean_degree}-seed{seed}")

    @classmethod
    def batch(cls, kind, n, exponent, mean_degree, seed, relabel: bool = True) -> "HostGraph":
        """Block-diagonal union of many ER (kind 0) / Chung-Lu (kind 1)
        graphs; ``.voff`` holds the vertex offsets (graph g on
        [voff[g], voff[g+1]))."""
        kind = np.ascontiguousarray(kind, np.int32)
        ng = len(kind)
        n = np.ascontiguousarray(n, np.int64)
        exponent = np.ascontiguousarray(exponent, np.float64)
        mean_degree = np.ascontiguousarray(mean_degree, np.float64)
        seed = np.ascontiguousarray(seed, np.uint64)
        voff = np.zeros(ng + 1, np.int64)
        h = C.c_void_p()
        check(lib().gk_hgraph_batch(ng, _ptr(kind), _ptr(n, C.c_int64), _ptr(exponent, C.c_double),
                                    _ptr(mean_degree, C.c_double), _ptr(seed, C.c_uint64), int(relabel), C.byref(h),
                                    _ptr(voff, C.c_int64)), "gk_hgraph_batch")
        g = cls(h, f"batch-{ng}")
        g.voff = voff
        return g

    @classmethod
    def grid(cls, rows: int, cols: int | None = None, diagonal: bool = True, relabel: bool = True) -> "HostGraph":
        h = C.c_void_p()
        cols = rows if cols is None else cols
        check(lib().gk_hgraph_grid(rows, cols, int(diagonal), int(relabel), C.byref(h)), "gk_hgraph_grid")
        return cls(h, f"grid-{rows}x{cols}{'-diag' if diagonal else ''}")

    @classmethod
    def from_edges(cls, n: int, u, v, relabel: bool = False) -> "HostGraph":
        u = np.ascontiguousarray(u, np.int32)
        v = np.ascontiguousarray(v, np.int32)
        h = C.c_void_p()
        check(lib().gk_hgraph_from_edges(n, len(u), _ptr(u), _ptr(v), int(relabel), C.byref(h)),
              "gk_hgraph_from_edges")
        return cls(h, f"edges-n{n}")

    @classmethod
    def from_spmtx(cls, a: SpMtx, relabel: bool = False) -> "HostGraph":
        h = C.c_void_p()
        csr = a.gk(with_values=False)
        check(lib().gk_hgraph_from_csr(C.byref(csr), int(relabel), C.byref(h)), "gk_hgraph_from_csr")
        return cls(h, "csr")


def batch_config(ngraphs: int = 4096, seed: int = 1, n_min: int = 1024, n_max: int = 65536,
                 cl_mean_degree: float = 16.0) -> dict:
    """Seeded description of a batch of uneven graphs: vertex counts
    log-uniform in [n_min, n_max]; even-numbered graphs Erdos-Renyi with
    mean degree uniform in [8, 32], odd-numbered Chung-Lu with exponent
    uniform in [2.1, 2.8] and mean degree ``cl_mean_degree``."""
    rng = np.random.default_rng(seed)
    n = np.exp(rng.uniform(np.log(n_min), np.log(n_max), ngraphs)).astype(np.int64)
    kind = (np.arange(ngraphs) % 2).astype(np.int32)
    exponent = np.where(kind == 1, rng.uniform(2.1, 2.8, ngraphs), 0.0)
    mean_degree = np.where(kind == 0, rng.uniform(8.0, 32.0, ngraphs), cl_mean_degree)
    seeds = rng.integers(1, 2 ** 62, ngraphs).astype(np.uint64)
    return {"kind": kind, "n": n, "exponent": exponent, "mean_degree": mean_degree, "seed": seeds}

