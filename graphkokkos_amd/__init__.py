"""graphkokkos_amd: GraphKokkos's masked sparse-sparse product and its
triangle-counting, k-truss and batched betweenness drivers on one AMD
Instinct MI355X (gfx950), behind the C ABI of libgraphhip.so."""
from .graph import (SpMtx, HostGraph, build_adjacency_matrix, build_symm_from_lower,
                    extract_lower_triangle, transpose, write_matrix)
from .algorithms import (Context, DeviceCsr, DeviceBuffer, default_context, mxmm_msa, mxmm_mca, mxmm_msa_cmask,
                         triangle_counting_masked_sandia, triangle_counting_vertex, triangle_counting_masked_lu,
                         k_truss, betweenness_centrality_batch)
from ._lib import (ACC_MSA_LDS, ACC_MCA_WAVE, ACC_HASH_GLOBAL, ACC_DENSE_GLOBAL, ACC_NAMES,
                   SEL_HEURISTIC, SEL_LEARNED, SEL_FORCED, GraphHipError, default_opts)

__all__ = [
    "SpMtx", "HostGraph", "build_adjacency_matrix", "build_symm_from_lower", "extract_lower_triangle",
    "transpose", "Context", "DeviceCsr", "DeviceBuffer", "default_context", "mxmm_msa", "mxmm_mca", "mxmm_msa_cmask",
    "triangle_counting_masked_sandia", "triangle_counting_vertex", "triangle_counting_masked_lu", "k_truss", "betweenness_centrality_batch", "ACC_MSA_LDS",
    "ACC_MCA_WAVE", "ACC_HASH_GLOBAL", "ACC_DENSE_GLOBAL", "ACC_NAMES", "SEL_HEURISTIC", "SEL_LEARNED",
    "SEL_FORCED", "GraphHipError", "default_opts",
]
