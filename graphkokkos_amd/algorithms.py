"""GraphKokkos's masked-product entry points and drivers, on the GPU.

Names, argument meaning and error behaviour follow the reference:

* ``mxmm_msa`` / ``mxmm_mca`` (include/matrix_la.h:87-123, 478-491, 562-574):
  ``C<M> = A*B``; C takes M's pattern, C.Val gets the products, explicit
  zeros kept.  Both names run the same HIP engine (the accumulator is picked
  per row); ``isParallel`` is accepted for signature compatibility.
* ``triangle_counting_masked_sandia(A, matrixMult, isParallel)``
  (src/main.cpp:69-90): ``sum(C<L> = L*L)`` with L the strict lower part.
* ``k_truss(A, k, matrixMult, isParallel)`` (src/main.cpp:93-141): repeats
  ``S<C> = C*C``, keeps ``S >= k-2``, until nothing is removed; returns the
  pattern (values 1).
* ``betweenness_centrality_batch(isParallel, A, batchSize)``
  (include/betweenness_centrality.h:406-506): sources ``0..batch-1``, result
  as float32 like the reference; ``bc_batch_f64`` takes an explicit source
  list and returns FP64.

Every compute call goes through libgraphhip.so; a missing library raises.
"""
from __future__ import annotations

import ctypes as C
from typing import Callable

import numpy as np

from . import _lib
from ._lib import GkCsr, GkOpts, GkTiming, check, default_opts, lib
from .graph import SpMtx, _ptr, extract_lower_triangle, transpose


class DeviceCsr:
    """A CSR resident in HBM (gk_dcsr)."""

    def __init__(self, ctx: "Context", a: SpMtx):
        self.ctx = ctx
        self.m, self.n, self.nz = a.m, a.n, a.nz
        h = C.c_void_p()
        csr = a.gk()
        check(lib().gk_dcsr_upload(ctx.handle, C.byref(csr), C.byref(h)), "gk_dcsr_upload")
        self.handle = h

    @classmethod
    def adopt(cls, ctx: "Context", handle: C.c_void_p) -> "DeviceCsr":
        """Wrap a gk_dcsr the library built (freed with this object)."""
        obj = cls.__new__(cls)
        obj.ctx, obj.handle = ctx, handle
        m, n, nz = C.c_int64(), C.c_int64(), C.c_int64()
        check(lib().gk_dcsr_dims(handle, C.byref(m), C.byref(n), C.byref(nz)), "gk_dcsr_dims")
        obj.m, obj.n, obj.nz = int(m.value), int(n.value), int(nz.value)
        return obj

    def to_host(self, values: bool = False) -> SpMtx:
        """Rows and columns; the values too when `values` (a library-built
        CSR that holds them, such as a complement-mask product)."""
        rst = np.zeros(self.m + 1, np.int32)
        col = np.zeros(max(self.nz, 1), np.int32)
        check(lib().gk_dcsr_download(self.ctx.handle, self.handle, _ptr(rst), _ptr(col)), "gk_dcsr_download")
        val = None
        if values:
            val = np.zeros(max(self.nz, 1), np.int32)
            check(lib().gk_dcsr_download_val(self.ctx.handle, self.handle, _ptr(val)), "gk_dcsr_download_val")
            val = val[:self.nz].copy()
        return SpMtx(self.m, self.n, rst, col[:self.nz].copy(), val)

    def free(self):
        if self.handle:
            lib().gk_dcsr_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceBuffer:
    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx, self.nbytes = ctx, nbytes
        p = C.c_void_p()
        check(lib().gk_dmalloc(ctx.handle, nbytes, C.byref(p)), "gk_dmalloc")
        self.ptr = p

    def to_host(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype)
        if count:
            check(lib().gk_memcpy_d2h(self.ctx.handle, out.ctypes.data_as(C.c_void_p), self.ptr,
                                      out.nbytes), "gk_memcpy_d2h")
        return out

    def free(self):
        if self.ptr:
            lib().gk_dfree(self.ctx.handle, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """gk_ctx: one device, one stream (counterpart of Kokkos::initialize)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().gk_init(device, C.byref(h)), "gk_init")
        self.handle = h

    def close(self):
        if self.handle:
            lib().gk_finalize(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def sync(self):
        check(lib().gk_sync(self.handle), "gk_sync")

    def upload(self, a: SpMtx) -> DeviceCsr:
        return DeviceCsr(self, a)

    def alloc(self, nbytes: int) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    # ---- on-device graph construction (the HostGraph generators on the GPU) ----
    def _dgraph(self, fn: str, *args, relabel: bool = True, perm: bool = False):
        A, L = C.c_void_p(), C.c_void_p()
        buf = None
        if perm:
            buf = self.alloc(4 * max(int(args[-1]), 1))
            args = args[:-1]
        check(getattr(lib(), fn)(self.handle, *args, int(relabel), C.byref(A), C.byref(L),
                                 buf.ptr if buf is not None else None), fn)
        out = (DeviceCsr.adopt(self, A), DeviceCsr.adopt(self, L))
        if perm:
            p = buf.to_host(np.int32, out[0].m)
            buf.free()
            return out + (p,)
        return out

    def dgraph_rmat(self, scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
                    seed: int = 1, relabel: bool = True, perm: bool = False):
        """R-MAT built in HBM: (A, L) as DeviceCsr (and the relabelling with perm=True)."""
        extra = (1 << scale,) if perm else ()
        return self._dgraph("gk_dgraph_rmat", scale, edge_factor, a, b, c, seed, *extra, relabel=relabel, perm=perm)

    def dgraph_er(self, n: int, mean_degree: float, seed: int = 1, relabel: bool = True, perm: bool = False):
        extra = (n,) if perm else ()
        return self._dgraph("gk_dgraph_er", n, mean_degree, seed, *extra, relabel=relabel, perm=perm)

    def dgraph_chung_lu(self, n: int, exponent: float, mean_degree: float = 16.0, seed: int = 1,
                        relabel: bool = True, perm: bool = False):
        extra = (n,) if perm else ()
        return self._dgraph("gk_dgraph_chung_lu", n, exponent, mean_degree, seed, *extra, relabel=relabel,
                            perm=perm)

    def dgraph_grid(self, rows: int, cols: int | None = None, diagonal: bool = True, relabel: bool = True,
                    perm: bool = False):
        cols = rows if cols is None else cols
        extra = (rows * cols,) if perm else ()
        return self._dgraph("gk_dgraph_grid", rows, cols, int(diagonal), *extra, relabel=relabel, perm=perm)

    def last_timing(self) -> dict:
        t = GkTiming()
        check(lib().gk_last_timing(self.handle, C.byref(t)), "gk_last_timing")
        return {"total_ms": t.total_ms, "kernel_ms": t.kernel_ms, "select_ms": t.select_ms,
                "launches": t.launches, "iterations": t.iterations,
                "rows_per_acc": list(t.rows_per_acc), "acc_ms": list(t.acc_ms), "acc_bytes": list(t.acc_bytes),
                "acc_launches": list(t.acc_launches), "bytes_alg": t.bytes_alg, "flop": t.flop,
                "hits": t.hits}

    @staticmethod
    def _opts(opts: GkOpts | dict | None) -> GkOpts:
        if opts is None:
            return default_opts()
        if isinstance(opts, dict):
            return default_opts(**opts)
        return opts

    # ---- products -------------------------------------------------------
    def masked_spgemm(self, A: SpMtx, B: SpMtx, M: SpMtx, opts=None) -> np.ndarray:
        out = np.zeros(M.nz, np.int32)
        a, b, m = A.gk(), B.gk(), M.gk(with_values=False)
        o = self._opts(opts)
        check(lib().gk_masked_spgemm_i32(self.handle, C.byref(a), C.byref(b), C.byref(m), _ptr(out),
                                         C.byref(o)), "gk_masked_spgemm_i32")
        return out

    def masked_spgemm_dev(self, A: DeviceCsr, B: DeviceCsr, M: DeviceCsr, out: DeviceBuffer, opts=None) -> None:
        o = self._opts(opts)
        check(lib().gk_masked_spgemm_dev(self.handle, A.handle, B.handle, M.handle, out.ptr, C.byref(o)),
              "gk_masked_spgemm_dev")

    def masked_spgemm_cmask_dev(self, A: DeviceCsr, B: DeviceCsr, M: DeviceCsr, dense_rows: int = -1,
                                out: DeviceCsr | None = None) -> DeviceCsr:
        """C<not M> = A*B on device handles (M with sorted rows); the result
        is a library-built CSR with values: a new one, or `out` (a result of
        an earlier call) refilled, its storage grown only when too small.
        dense_rows: -1 picks each row's way (sorted or dense accumulation),
        0 / 1 force one, 2 forces dense rows counted first (no work pool)."""
        h = C.c_void_p(out.handle.value if out is not None else None)
        check(lib().gk_masked_spgemm_cmask_dev(self.handle, A.handle, B.handle, M.handle, dense_rows, C.byref(h)),
              "gk_masked_spgemm_cmask_dev")
        if out is None:
            return DeviceCsr.adopt(self, h)
        m, n, nz = C.c_int64(), C.c_int64(), C.c_int64()
        check(lib().gk_dcsr_dims(h, C.byref(m), C.byref(n), C.byref(nz)), "gk_dcsr_dims")
        out.m, out.n, out.nz = int(m.value), int(n.value), int(nz.value)
        return out

    def masked_spgemm_cmask(self, A: SpMtx, B: SpMtx, M: SpMtx, dense_rows: int = -1) -> SpMtx:
        """C<not M> = A*B: the entries some product term reaches outside M's
        pattern, with their int32 sums (explicit zeros kept)."""
        Ms, _ = _canonical_mask(M)
        dA, dB, dM = self.upload(A), self.upload(B), self.upload(Ms)
        try:
            dC = self.masked_spgemm_cmask_dev(dA, dB, dM, dense_rows)
            try:
                return dC.to_host(values=True)
            finally:
                dC.free()
        finally:
            for d in (dA, dB, dM):
                d.free()

    def tc_vertex(self, A: SpMtx, opts=None) -> np.ndarray:
        """(row sums of SQ<A> = A*A) >> 1 in int32: triangles through each
        vertex of a symmetric pattern A (A's rows sorted)."""
        out = np.zeros(max(A.m, 1), np.int32)
        a = A.gk()
        o = self._opts(opts)
        check(lib().gk_tc_vertex(self.handle, C.byref(a), _ptr(out), C.byref(o)), "gk_tc_vertex")
        return out[:A.m]

    def tc_count(self, L: SpMtx | DeviceCsr, opts=None) -> int:
        o = self._opts(opts)
        cnt = C.c_uint64(0)
        if isinstance(L, DeviceCsr):
            check(lib().gk_tc_count_dev(self.handle, L.handle, C.byref(cnt), C.byref(o)), "gk_tc_count_dev")
        else:
            csr = L.gk(with_values=False)
            check(lib().gk_tc_count(self.handle, C.byref(csr), C.byref(cnt), C.byref(o)), "gk_tc_count")
        return int(cnt.value)

    def tc_rows(self, L: DeviceCsr, opts=None) -> np.ndarray:
        o = self._opts(opts)
        buf = self.alloc(8 * max(L.m, 1))
        check(lib().gk_tc_rows_dev(self.handle, L.handle, buf.ptr, C.byref(o)), "gk_tc_rows_dev")
        return buf.to_host(np.uint64, L.m)

    def ktruss_support(self, A: SpMtx, opts=None) -> np.ndarray:
        o = self._opts(opts)
        out = np.zeros(A.nz, np.uint32)
        csr = A.gk(with_values=False)
        check(lib().gk_ktruss_support(self.handle, C.byref(csr), _ptr(out, C.c_uint32), C.byref(o)),
              "gk_ktruss_support")
        return out

    def ktruss(self, A: SpMtx, k: int, opts=None) -> tuple[SpMtx, int]:
        o = self._opts(opts)
        rst = np.zeros(A.m + 1, np.int32)
        col = np.zeros(max(A.nz, 1), np.int32)
        nz = C.c_int64(0)
        it = C.c_int64(0)
        csr = A.gk(with_values=False)
        check(lib().gk_ktruss(self.handle, C.byref(csr), k, _ptr(rst), _ptr(col), C.byref(nz), C.byref(it),
                              C.byref(o)), "gk_ktruss")
        res = SpMtx(A.m, A.n, rst, col[:nz.value].copy(), np.ones(nz.value, np.int32), A.matcode)
        return res, int(it.value)

    def truss_decompose(self, A: SpMtx | DeviceCsr, opts=None) -> tuple[np.ndarray, int]:
        o = self._opts(opts)
        kmax = C.c_int32(0)
        if isinstance(A, DeviceCsr):
            buf = self.alloc(4 * max(A.nz, 1))
            check(lib().gk_truss_decompose_dev(self.handle, A.handle, buf.ptr, C.byref(kmax), C.byref(o)),
                  "gk_truss_decompose_dev")
            return buf.to_host(np.uint32, A.nz), int(kmax.value)
        out = np.zeros(A.nz, np.uint32)
        csr = A.gk(with_values=False)
        check(lib().gk_truss_decompose(self.handle, C.byref(csr), _ptr(out, C.c_uint32), C.byref(kmax),
                                       C.byref(o)), "gk_truss_decompose")
        return out, int(kmax.value)

    def bc_batch_f64(self, A: SpMtx, sources, opts=None) -> np.ndarray:
        o = self._opts(opts)
        src = np.ascontiguousarray(sources, np.int32)
        out = np.zeros(A.m, np.float64)
        csr = A.gk(with_values=False)
        check(lib().gk_bc_batch_f64(self.handle, C.byref(csr), _ptr(src), len(src), _ptr(out, C.c_double),
                                    C.byref(o)), "gk_bc_batch_f64")
        return out

    def batch_tc_ktruss(self, A: DeviceCsr, voff, k: int = 4, use_graph: bool = True, reps: int = 1,
                        keep: bool = False, digests: bool = False, opts=None) -> dict:
        """Triangle count and k-truss of every graph of a block-diagonal union
        (gk_batch_tc_ktruss_dev)."""
        o = self._opts(opts)
        voff = np.ascontiguousarray(voff, np.int64)
        ng = len(voff) - 1
        tri = np.zeros(ng, np.uint64)
        te = np.zeros(ng, np.int64)
        ms = C.c_double(0)
        dig = np.zeros(reps, np.uint64) if digests else None
        kb = self.alloc(max(A.nz, 1)) if keep else None
        check(lib().gk_batch_tc_ktruss_dev(self.handle, A.handle, ng, _ptr(voff, C.c_int64), k, int(use_graph), reps,
                                           _ptr(tri, C.c_uint64), _ptr(te, C.c_int64), kb.ptr if kb else None,
                                           C.byref(ms), _ptr(dig, C.c_uint64) if dig is not None else None,
                                           C.byref(o)), "gk_batch_tc_ktruss_dev")
        out = {"triangles": tri, "truss_edges": te, "ms": ms.value}
        if dig is not None:
            out["digests"] = dig
        if kb is not None:
            out["keep"] = kb.to_host(np.uint8, A.nz).astype(bool)
            kb.free()
        return out

    def bc_batch_dev(self, A: DeviceCsr, sources, out: DeviceBuffer, AT: DeviceCsr | None = None,
                     opts=None) -> None:
        o = self._opts(opts)
        src = np.ascontiguousarray(sources, np.int32)
        check(lib().gk_bc_batch_dev(self.handle, A.handle, AT.handle if AT else None, _ptr(src), len(src),
                                    out.ptr, C.byref(o)), "gk_bc_batch_dev")

    def bc_batch_dev_noalloc(self, A: DeviceCsr, sources, out: DeviceBuffer, AT: DeviceCsr | None = None) -> None:
        """gk_bc_batch_dev_noalloc: the same scores with every work buffer in
        the library's fixed per-device scratch array (no device allocation)."""
        src = np.ascontiguousarray(sources, np.int32)
        check(lib().gk_bc_batch_dev_noalloc(self.handle, A.handle, AT.handle if AT else None, _ptr(src), len(src),
                                            out.ptr), "gk_bc_batch_dev_noalloc")


_default: Context | None = None


def default_context() -> Context:
    global _default
    if _default is None:
        _default = Context(0)
    return _default


# ---- reference-named entry points ------------------------------------------

def _canonical_mask(M: SpMtx):
    """(M', pos): M' holds M's pattern with every row sorted and duplicate
    free, pos[p] is the entry of M' that entry p of M names; (M, None) when
    M is already canonical.  The reference's MSA reads its accumulator once
    per mask entry (include/matrix_la.h:612-631), so rows in file order and
    repeated columns are legal masks and every copy gets the column's value."""
    rows = np.repeat(np.arange(M.m, dtype=np.int64), np.diff(M.Rst))
    same = rows[1:] == rows[:-1]
    if M.nz < 2 or bool(np.all(M.Col[1:][same] > M.Col[:-1][same])):
        return M, None
    key = rows * max(M.n, 1) + M.Col.astype(np.int64)
    uniq, pos = np.unique(key, return_inverse=True)
    urow = uniq // max(M.n, 1)
    rst = np.zeros(M.m + 1, np.int64)
    np.cumsum(np.bincount(urow, minlength=M.m), out=rst[1:])
    Ms = SpMtx(M.m, M.n, rst.astype(np.int32), (uniq % max(M.n, 1)).astype(np.int32), None, M.matcode)
    return Ms, pos


def _product_into(A: SpMtx, B: SpMtx, M: SpMtx, C_: SpMtx) -> None:
    if A.n != B.m:
        raise ValueError("A.n must equal B.m")
    Ms, pos = _canonical_mask(M)
    vals = default_context().masked_spgemm(A, B, Ms)
    if pos is not None:
        vals = vals[pos]
    # C.resizeRows(M.m); C.resizeVals(M.nz); C.n = M.n; copy M's pattern
    C_.m, C_.n = M.m, M.n
    C_.Rst = M.Rst.copy()
    C_.Col = M.Col.copy()
    C_.Val = vals
    C_.matcode = M.matcode


def mxmm_msa(isParallel: bool, A: SpMtx, B: SpMtx, M: SpMtx, C_: SpMtx) -> None:
    """C<M> = A*B (include/matrix_la.h:562-574)."""
    _product_into(A, B, M, C_)


def mxmm_mca(isParallel: bool, A: SpMtx, B: SpMtx, M: SpMtx, C_: SpMtx) -> None:
    """C<M> = A*B (include/matrix_la.h:478-491)."""
    _product_into(A, B, M, C_)


def mxmm_msa_cmask(isParallel: bool, A: SpMtx, B: SpMtx, M: SpMtx, C_: SpMtx) -> None:
    """C<not M> = A*B (include/matrix_la.h:733-934): C gets M's shape and
    the entries some term A(i,k)*B(k,j) reaches outside M's pattern, rows
    in column order, explicit zeros kept.  Repeated mask columns count once
    (the reference's symbolic pass would subtract them twice)."""
    if A.n != B.m:
        raise ValueError("A.n must equal B.m")
    R = default_context().masked_spgemm_cmask(A, B, M)
    C_.m, C_.n = R.m, R.n
    C_.Rst, C_.Col, C_.Val = R.Rst, R.Col, R.Val


MxmOp = Callable[[bool, SpMtx, SpMtx, SpMtx, SpMtx], None]


def triangle_counting_masked_sandia(A: SpMtx, matrixMult: MxmOp = mxmm_msa, isParallel: bool = True) -> int:
    """sum(C<L> = L*L), L = extract_lower_triangle(A) (src/main.cpp:69-90)."""
    L = extract_lower_triangle(A)
    if matrixMult in (mxmm_msa, mxmm_mca):
        return default_context().tc_count(L)
    C_ = SpMtx()
    matrixMult(isParallel, L, L, L, C_)
    return int(np.asarray(C_.Val, np.int64).sum())


def _rows_sorted(A: SpMtx) -> bool:
    rows = np.repeat(np.arange(A.m), np.diff(A.Rst))
    same = rows[1:] == rows[:-1]
    return bool(np.all(A.Col[1:][same] > A.Col[:-1][same]))


def triangle_counting_vertex(A: SpMtx, matrixMult: MxmOp = mxmm_msa, isParallel: bool = True) -> np.ndarray:
    """Triangles through every vertex: (row sums of SQ<A> = A*A) >> 1 in
    int32 arithmetic (src/main.cpp:16-39).  With this library's product and
    sorted rows the sums are taken on the device (gk_tc_vertex)."""
    if matrixMult in (mxmm_msa, mxmm_mca) and A.m == A.n and _rows_sorted(A):
        return default_context().tc_vertex(A)
    SQ = SpMtx()
    matrixMult(isParallel, A, A, A, SQ)
    rows = np.repeat(np.arange(SQ.m), np.diff(SQ.Rst))
    sums = np.zeros(SQ.m, np.int64)
    np.add.at(sums, rows, np.asarray(SQ.Val, np.int64))
    return ((sums & 0xFFFFFFFF).astype(np.uint32).view(np.int32) >> 1).astype(np.int32)


def triangle_counting_masked_lu(A: SpMtx, matrixMult: MxmOp = mxmm_msa, isParallel: bool = True) -> int:
    """sum(C<A> = L*U) / 2 with L the strictly lower part of A and U its
    transpose (src/main.cpp:42-66)."""
    L = extract_lower_triangle(A)
    U = transpose(L)
    C_ = SpMtx()
    matrixMult(isParallel, L, U, A, C_)
    return int(np.asarray(C_.Val, np.int64).sum()) >> 1


def k_truss(A: SpMtx, k: int, matrixMult: MxmOp = mxmm_msa, isParallel: bool = True) -> SpMtx:
    """k-truss pattern of the symmetric A (src/main.cpp:93-141)."""
    if matrixMult in (mxmm_msa, mxmm_mca):
        res, _ = default_context().ktruss(A, k)
        return res
    # An arbitrary product: the reference's loop, compaction on the host.
    Cm = A
    for _ in range(max(A.m, 1)):
        T = SpMtx()
        matrixMult(isParallel, Cm, Cm, Cm, T)
        keep = np.asarray(T.Val) >= k - 2
        rows = np.repeat(np.arange(Cm.m), np.diff(Cm.Rst))
        rst = np.zeros(Cm.m + 1, np.int64)
        np.cumsum(np.bincount(rows[keep], minlength=Cm.m), out=rst[1:])
        nxt = SpMtx(Cm.m, Cm.n, rst.astype(np.int32), Cm.Col[keep], np.ones(int(keep.sum()), np.int32),
                    Cm.matcode)
        if nxt.nz == Cm.nz:
            return nxt
        Cm = nxt
    return Cm


def betweenness_centrality_batch(isParallel: bool, A: SpMtx, batchSize: int) -> np.ndarray:
    """BC over sources 0..min(n, batchSize)-1, float32 result
    (include/betweenness_centrality.h:406-506)."""
    if A.m != A.n:
        raise ValueError("non-square matrix; BC is only for square matrices")
    n = min(A.n, int(batchSize))
    return default_context().bc_batch_f64(A, np.arange(n, dtype=np.int32)).astype(np.float32)
