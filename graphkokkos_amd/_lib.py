"""ctypes binding of libgraphhip.so (declared in include/graphhip.h).

The library is built in-tree by ``python -m graphkokkos_amd.build``.  There is
no fallback: if the shared library is missing, importing the product API
raises, and every compute entry point runs the HIP kernels.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libgraphhip.so"

GK_OK = 0
GK_EINVAL, GK_ENOMEM, GK_EHIP, GK_EIO, GK_ERANGE, GK_ESTATE = -1, -2, -3, -4, -5, -6

ACC_AUTO, ACC_MSA_LDS, ACC_MCA_WAVE, ACC_HASH_GLOBAL, ACC_DENSE_GLOBAL = -1, 0, 1, 2, 3
NUM_ACC = 4
NUM_KERNELS = 6
ACC_NAMES = ["msa-lds", "mca-wave", "hash-global", "dense-global"]
KERNEL_NAMES = ["msa-lds", "mca-wave", "hash-global", "dense-global", "msa-lds-long", "peel-stream"]
SEL_HEURISTIC, SEL_LEARNED, SEL_FORCED = 0, 1, 2

i32p = C.POINTER(C.c_int32)
u32p = C.POINTER(C.c_uint32)
i64p = C.POINTER(C.c_int64)
u64p = C.POINTER(C.c_uint64)
f64p = C.POINTER(C.c_double)


class GkCsr(C.Structure):
    _fields_ = [("m", C.c_int64), ("n", C.c_int64), ("nz", C.c_int64),
                ("rst", i32p), ("col", i32p), ("val", i32p)]


class GkOpts(C.Structure):
    _fields_ = [("struct_size", C.c_int32), ("selector", C.c_int32), ("forced_accum", C.c_int32),
                ("mca_max_mask", C.c_int32), ("lds_max_mask", C.c_int32), ("dense_min_mask", C.c_int32),
                ("split_flop", C.c_int32), ("block_threads", C.c_int32), ("verbose", C.c_int32),
                ("small_max_mask", C.c_int32), ("blocks_small", C.c_int32), ("blocks_long", C.c_int32),
                ("blocks_mca", C.c_int32), ("blocks_global", C.c_int32), ("batch_blocks", C.c_int32),
                ("batch_order", C.c_int32), ("hub_window", C.c_int32), ("learned_margin", C.c_int32),
                ("hub_probe", C.c_int32), ("selector_labels", C.c_int32)]


class GkTiming(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("kernel_ms", C.c_double), ("select_ms", C.c_double),
                ("launches", C.c_int64), ("iterations", C.c_int64), ("rows_per_acc", C.c_int64 * NUM_ACC),
                ("acc_ms", C.c_double * NUM_KERNELS), ("acc_bytes", C.c_double * NUM_KERNELS),
                ("acc_launches", C.c_int64 * NUM_KERNELS), ("bytes_alg", C.c_double), ("flop", C.c_double),
                ("hits", C.c_double)]


class GraphHipError(RuntimeError):
    def __init__(self, rc: int, where: str, msg: str):
        super().__init__(f"{where} failed ({rc}): {msg}")
        self.rc = rc


_SIGS = {
    "gk_last_error": (C.c_char_p, []),
    "gk_abi_version": (C.c_int, []),
    "gk_init": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "gk_finalize": (None, [C.c_void_p]),
    "gk_sync": (C.c_int, [C.c_void_p]),
    "gk_set_host_threads": (None, [C.c_int]),
    "gk_opts_default": (None, [C.POINTER(GkOpts)]),
    "gk_last_timing": (C.c_int, [C.c_void_p, C.POINTER(GkTiming)]),
    "gk_hgraph_rmat": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_uint64, C.c_int,
                                 C.POINTER(C.c_void_p)]),
    "gk_hgraph_er": (C.c_int, [C.c_int64, C.c_double, C.c_uint64, C.c_int, C.POINTER(C.c_void_p)]),
    "gk_hgraph_chung_lu": (C.c_int, [C.c_int64, C.c_double, C.c_double, C.c_uint64, C.c_int,
                                     C.POINTER(C.c_void_p)]),
    "gk_hgraph_grid": (C.c_int, [C.c_int64, C.c_int64, C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "gk_hgraph_batch": (C.c_int, [C.c_int64, i32p, i64p, f64p, f64p, u64p, C.c_int, C.POINTER(C.c_void_p), i64p]),
    "gk_hgraph_from_edges": (C.c_int, [C.c_int64, C.c_int64, i32p, i32p, C.c_int, C.POINTER(C.c_void_p)]),
    "gk_hgraph_from_csr": (C.c_int, [C.POINTER(GkCsr), C.c_int, C.POINTER(C.c_void_p)]),
    "gk_hgraph_view": (C.c_int, [C.c_void_p, C.POINTER(GkCsr), C.POINTER(GkCsr)]),
    "gk_hgraph_perm": (C.c_int, [C.c_void_p, i32p]),
    "gk_hgraph_free": (None, [C.c_void_p]),
    "gk_mtx_read": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "gk_mtx_view": (C.c_int, [C.c_void_p, C.POINTER(GkCsr), C.POINTER(f64p), C.c_char_p]),
    "gk_mtx_free": (None, [C.c_void_p]),
    "gk_dcsr_upload": (C.c_int, [C.c_void_p, C.POINTER(GkCsr), C.POINTER(C.c_void_p)]),
    "gk_dcsr_free": (None, [C.c_void_p]),
    "gk_dcsr_nnz": (C.c_int64, [C.c_void_p]),
    "gk_matrix_read": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)]),
    "gk_matrix_write": (C.c_int, [C.c_char_p, C.c_char_p, C.c_void_p, C.POINTER(C.c_int32), C.c_char_p]),
    "gk_dcsr_dims": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "gk_dcsr_download": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "gk_dcsr_download_val": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_int32)]),
    "gk_masked_spgemm_cmask_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                             C.POINTER(C.c_void_p)]),
    "gk_dgraph_rmat": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                 C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_void_p]),
    "gk_dgraph_er": (C.c_int, [C.c_void_p, C.c_int64, C.c_double, C.c_uint64, C.c_int, C.POINTER(C.c_void_p),
                               C.POINTER(C.c_void_p), C.c_void_p]),
    "gk_dgraph_chung_lu": (C.c_int, [C.c_void_p, C.c_int64, C.c_double, C.c_double, C.c_uint64, C.c_int,
                                     C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_void_p]),
    "gk_dgraph_grid": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_int, C.c_int, C.POINTER(C.c_void_p),
                                 C.POINTER(C.c_void_p), C.c_void_p]),
    "gk_masked_spgemm_i32": (C.c_int, [C.c_void_p, C.POINTER(GkCsr), C.POINTER(GkCsr), C.POINTER(GkCsr), i32p,
                                       C.POINTER(GkOpts)]),
    "gk_masked_spgemm_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.POINTER(GkOpts)]),
    "gk_tc_count": (C.c_int, [C.c_void_p, C.POINTER(GkCsr), u64p, C.POINTER(GkOpts)]),
    "gk_tc_vertex": (C.c_int, [C.c_void_p, C.POINTER(GkCsr), i32p, C.POINTER(GkOpts)]),
    "gk_tc_vertex_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(GkOpts)]),
    "gk_tc_count_dev": (C.c_int, [C.c_void_p, C.c_void_p, u64p, C.POINTER(GkOpts)]),
    "gk_tc_rows_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(GkOpts)]),
    "gk_ktruss_support": (C.c_int, [C.c_void_p, C.POINTER(GkCsr), u32p, C.POINTER(GkOpts)]),
    "gk_ktruss": (C.c_int, [C.c_void_p, C.POINTER(GkCsr), C.c_int, i32p, i32p, i64p, i64p, C.POINTER(GkOpts)]),
    "gk_truss_decompose": (C.c_int, [C.c_void_p, C.POINTER(GkCsr), u32p, i32p, C.POINTER(GkOpts)]),
    "gk_truss_decompose_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, i32p, C.POINTER(GkOpts)]),
    "gk_bc_batch_f64": (C.c_int, [C.c_void_p, C.POINTER(GkCsr), i32p, C.c_int64, f64p, C.POINTER(GkOpts)]),
    "gk_bc_batch_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, i32p, C.c_int64, C.c_void_p,
                                  C.POINTER(GkOpts)]),
    "gk_bc_batch_dev_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, i32p, C.c_int64, C.c_void_p]),
    "gk_batch_tc_ktruss_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, i64p, C.c_int, C.c_int, C.c_int, u64p,
                                         i64p, C.c_void_p, f64p, u64p, C.POINTER(GkOpts)]),
    "gk_model_num_params": (C.c_int, []),
    "gk_model_set": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int]),
    "gk_model_get": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int]),
    "gk_model_clear": (C.c_int, [C.c_void_p]),
    "gk_model_forward": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int64, C.POINTER(C.c_float)]),
    "gk_model_grad": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_uint8),
                                C.c_int64, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "gk_model_train": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_uint8),
                                 C.c_int64, C.c_int, C.c_int, C.c_float, C.c_uint64, C.POINTER(C.c_float)]),
    "gk_selector_samples": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                      C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_uint8),
                                      C.c_void_p]),
    "gk_dmalloc": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)]),
    "gk_dfree": (None, [C.c_void_p, C.c_void_p]),
    "gk_memcpy_d2h": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    "gk_memcpy_h2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    "gk_memcpy_d2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
}

_lib: C.CDLL | None = None


def lib() -> C.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m graphkokkos_amd.build` "
                               "(there is no CPU fallback)")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols() -> list[str]:
    return sorted(_SIGS)


def check(rc: int, where: str) -> None:
    if rc != GK_OK:
        msg = lib().gk_last_error().decode(errors="replace")
        raise GraphHipError(rc, where, msg)


def default_opts(**kw) -> GkOpts:
    o = GkOpts()
    lib().gk_opts_default(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(f"unknown option {k}")
        setattr(o, k, v)
    return o
