// Triangles through every vertex (GraphKokkos's triangle_counting_vertex,
// src/main.cpp:16-39): SQ<A> = A*A on A's pattern, then per row the int sum
// of SQ's values shifted right once.  The product runs on the masked-product
// engine into HBM; one wave per row sums it there, so SQ never reaches the
// host.
#include "gk_device.h"

using namespace gk;

namespace gk {
namespace {

// counts[i] = (int32)(sum of c over row i, wrapping) >> 1, one wave per row.
__global__ void k_row_halfsum(int64_t m, const int32_t *__restrict__ rst, const int32_t *__restrict__ c,
                              int32_t *__restrict__ counts) {
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < m; i += waves) {
        uint32_t s = 0;
        for (int32_t p = rst[i] + lane; p < rst[i + 1]; p += 64) s += (uint32_t)c[p];
        s = dev::wave_sum(s);
        if (lane == 0) counts[i] = (int32_t)s >> 1;
    }
}

}  // namespace
}  // namespace gk

extern "C" int gk_tc_vertex_dev(gk_ctx *ctx, const gk_dcsr *A, int32_t *d_counts, const gk_opts *opts) {
    GK_CHECK_ARG(ctx && A && (d_counts || A->m == 0), "null argument");
    GK_CHECK_ARG(A->m == A->n, "A must be square (%lld x %lld)", (long long)A->m, (long long)A->n);
    GK_TRY(require_sorted(A, "A"));
    int32_t *d_c = nullptr;
    if (hipMalloc(&d_c, sizeof(int32_t) * (size_t)(A->nz ? A->nz : 1)) != hipSuccess) {
        set_error("hipMalloc of %lld values for SQ failed", (long long)A->nz);
        return GK_ENOMEM;
    }
    int rc = gk_masked_spgemm_dev(ctx, A, A, A, d_c, opts);
    if (rc == GK_OK && A->m) {
        const int64_t blocks = std::min<int64_t>((A->m + 3) / 4, (int64_t)ctx->num_cus * 16);
        k_row_halfsum<<<(int)blocks, 256, 0, ctx->stream>>>(A->m, A->rst, d_c, d_counts);
        rc = launch_check("k_row_halfsum");
        if (rc == GK_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) {
            set_error("per-vertex triangle sums failed on the device");
            rc = GK_EHIP;
        }
    }
    (void)hipFree(d_c);
    return rc;
}

extern "C" int gk_tc_vertex(gk_ctx *ctx, const gk_csr *A, int32_t *counts, const gk_opts *opts) {
    GK_CHECK_ARG(ctx && A && (counts || A->m == 0), "null argument");
    GK_TRY(check_csr(A, "A"));
    gk_dcsr d;
    GK_TRY(upload_csr(ctx, A, &d));
    int32_t *d_counts = nullptr;
    if (hipMalloc(&d_counts, sizeof(int32_t) * (size_t)(A->m ? A->m : 1)) != hipSuccess) {
        free_csr(&d);
        set_error("hipMalloc of the vertex counts failed");
        return GK_ENOMEM;
    }
    int rc = gk_tc_vertex_dev(ctx, &d, d_counts, opts);
    if (rc == GK_OK && A->m && d2h(ctx, counts, d_counts, sizeof(int32_t) * (size_t)A->m) != hipSuccess) {
        set_error("copy of the vertex counts to host failed");
        rc = GK_EHIP;
    }
    (void)hipFree(d_counts);
    free_csr(&d);
    return rc;
}
