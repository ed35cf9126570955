// On-device graph construction: the seeded generators and the cleaning
// pipeline of host_graph.cpp (symmetrise, drop self loops, deduplicate,
// relabel by (degree descending, id ascending), sorted int32 CSR, strict
// lower part) run on the GPU and leave the graph in HBM.
//
// The generators are counter-based (edge e draws rng(seed, e, level)), so
// every edge is drawn independently, and the cleaning only sorts, so the
// result is the host builder's graph bit for bit.  The pipeline is five
// passes over 64-bit keys:
//   1. k_gen_*        one key (max(u,v) << 32 | min(u,v)) per generated
//                     edge, all ones for a dropped edge or a self loop;
//   2. radix sort + unique of the keys;
//   3. degrees, then the relabel as a sort of (~degree << 32 | id);
//   4. both directions of every edge in new labels, radix sorted: the rows
//      and their columns come out in order;
//   5. row pointers from the degrees, the strict lower part by a search per
//      row.
// The reference builds graphs on the host (build_adjacency_matrix,
// extract_lower_triangle: include/matrix.h:643-726); its R-MAT generator
// (include/matrix_utils.h) is unseeded, so the generators are this
// library's own, shared with host_graph.cpp.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>

#include <cmath>
#include <vector>

#include "gk_device.h"

using namespace gk;

namespace gk {
namespace {

constexpr uint64_t kNoEdge = ~0ull;

// The counter-based generator of gk_internal.h (SplitMix64 finaliser), on
// the device; the tests compare the two builders' graphs entry by entry.
__device__ __forceinline__ uint64_t d_mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ uint64_t d_rng(uint64_t seed, uint64_t stream, uint64_t ctr) {
    return d_mix64(d_mix64(seed ^ (stream * 0xd1b54a32d192ed03ull)) + ctr * 0x9e3779b97f4a7c15ull);
}
__device__ __forceinline__ double d_u01(uint64_t x) { return double(x >> 11) * (1.0 / 9007199254740992.0); }

__device__ __forceinline__ uint64_t edge_key(int64_t u, int64_t v) {
    if (u == v) return kNoEdge;
    const uint64_t hi = (uint64_t)(u > v ? u : v), lo = (uint64_t)(u > v ? v : u);
    return hi << 32 | lo;
}

// R-MAT: edge e descends `scale` quadrant levels (a, b, c, d).
__global__ void k_gen_rmat(int64_t ne, int scale, double a, double ab, double abc, uint64_t seed,
                           uint64_t *__restrict__ key) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t x = 0, y = 0;
        for (int l = 0; l < scale; ++l) {
            const double r = d_u01(d_rng(seed, (uint64_t)e, (uint64_t)l));
            const int q = r < a ? 0 : (r < ab ? 1 : (r < abc ? 2 : 3));
            x = (x << 1) | (q >> 1);
            y = (y << 1) | (q & 1);
        }
        key[e] = edge_key(x, y);
    }
}

// Erdos-Renyi G(n, m): both endpoints uniform.
__global__ void k_gen_er(int64_t ne, int64_t n, uint64_t seed, uint64_t *__restrict__ key) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t u = (int64_t)(d_u01(d_rng(seed, (uint64_t)e, 0)) * n);
        int64_t v = (int64_t)(d_u01(d_rng(seed, (uint64_t)e, 1)) * n);
        if (u >= n) u = n - 1;
        if (v >= n) v = n - 1;
        key[e] = edge_key(u, v);
    }
}

__device__ __forceinline__ int64_t cdf_draw(const double *__restrict__ cdf, int64_t n, uint64_t r) {
    const double x = d_u01(r);
    int64_t l = 0, h = n;  // first i with cdf[i] > x
    while (l < h) {
        const int64_t mid = l + ((h - l) >> 1);
        if (cdf[mid] <= x) l = mid + 1; else h = mid;
    }
    return l < n - 1 ? l : n - 1;
}

// Chung-Lu: both endpoints by inverse CDF of the weights (table from the host).
__global__ void k_gen_chung_lu(int64_t ne, int64_t n, const double *__restrict__ cdf, uint64_t seed,
                               uint64_t *__restrict__ key) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x)
        key[e] = edge_key(cdf_draw(cdf, n, d_rng(seed, (uint64_t)e, 0)), cdf_draw(cdf, n, d_rng(seed, (uint64_t)e, 1)));
}

// Grid: slots right, down (and down-right with the diagonal) of every cell.
__global__ void k_gen_grid(int64_t rows, int64_t cols, int per, uint64_t *__restrict__ key) {
    const int64_t ne = rows * cols * per;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t cell = e / per, kind = e % per, r = cell / cols, c = cell % cols;
        int64_t v = -1;
        if (kind == 0) { if (c + 1 < cols) v = cell + 1; }
        else if (kind == 1) { if (r + 1 < rows) v = cell + cols; }
        else if (r + 1 < rows && c + 1 < cols) v = cell + cols + 1;
        key[e] = v < 0 ? kNoEdge : edge_key(cell, v);
    }
}

__global__ void k_degrees(int64_t ne, const uint64_t *__restrict__ key, int32_t *__restrict__ deg) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
        atomicAdd(&deg[key[e] >> 32], 1);
        atomicAdd(&deg[key[e] & 0xffffffffull], 1);
    }
}

// (degree descending, id ascending) as one ascending 64-bit key.
__global__ void k_relabel_keys(int64_t n, const int32_t *__restrict__ deg, uint64_t *__restrict__ rk) {
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x)
        rk[v] = (uint64_t)(0x7fffffffu - (uint32_t)deg[v]) << 32 | (uint64_t)v;
}

__global__ void k_perm(int64_t n, const uint64_t *__restrict__ sorted_rk, int relabel, int32_t *__restrict__ perm) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        if (relabel) perm[sorted_rk[r] & 0xffffffffull] = (int32_t)r;
        else perm[r] = (int32_t)r;
    }
}

__global__ void k_new_degrees(int64_t n, const int32_t *__restrict__ deg, const int32_t *__restrict__ perm,
                              int32_t *__restrict__ ndeg) {
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x)
        ndeg[perm[v]] = deg[v];
    if (blockIdx.x == 0 && threadIdx.x == 0) ndeg[n] = 0;
}

__global__ void k_sym_keys(int64_t ne, const uint64_t *__restrict__ key, const int32_t *__restrict__ perm,
                           uint64_t *__restrict__ sk) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t a = (uint32_t)perm[key[e] >> 32], b = (uint32_t)perm[key[e] & 0xffffffffull];
        sk[2 * e] = a << 32 | b;
        sk[2 * e + 1] = b << 32 | a;
    }
}

__global__ void k_low32(int64_t nz, const uint64_t *__restrict__ sk, int32_t *__restrict__ col) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < nz; p += (int64_t)gridDim.x * blockDim.x)
        col[p] = (int32_t)(sk[p] & 0xffffffffull);
}

// Columns below the row: the leading run of a sorted row.
__global__ void k_lower_count(int64_t m, const int32_t *__restrict__ rst, const int32_t *__restrict__ col,
                              int32_t *__restrict__ len) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        int32_t l = rst[i], h = rst[i + 1];
        while (l < h) {
            const int32_t mid = l + ((h - l) >> 1);
            if (col[mid] < (int32_t)i) l = mid + 1; else h = mid;
        }
        len[i] = l - rst[i];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) len[m] = 0;
}

__global__ void k_lower_fill(int64_t m, const int32_t *__restrict__ rst, const int32_t *__restrict__ col,
                             const int32_t *__restrict__ lrst, int32_t *__restrict__ lcol) {
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < m; i += waves)
        for (int32_t x = lane; x < lrst[i + 1] - lrst[i]; x += 64) lcol[lrst[i] + x] = col[rst[i] + x];
}

// Device buffers of one build, freed when it returns.
struct Bufs {
    std::vector<void *> p;
    ~Bufs() {
        for (void *x : p) (void)hipFree(x);
    }
    template <class T>
    int alloc(T **out, int64_t n) {
        if (hipMalloc((void **)out, sizeof(T) * (size_t)(n > 0 ? n : 1)) != hipSuccess) {
            set_error("hipMalloc of %lld bytes for the graph build failed", (long long)(sizeof(T) * n));
            return GK_ENOMEM;
        }
        p.push_back(*out);
        return GK_OK;
    }
};

int grid_for(gk_ctx *ctx, int64_t n) {
    return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, (int64_t)ctx->num_cus * 16));
}

template <class KeyT>
int sort_keys(gk_ctx *ctx, Bufs &B, KeyT *in, KeyT *out, int64_t n, int end_bit) {
    size_t tmp = 0;
    GK_HIP(rocprim::radix_sort_keys(nullptr, tmp, in, out, (size_t)n, 0, end_bit, ctx->stream));
    void *t;
    GK_TRY(B.alloc((char **)&t, (int64_t)tmp));
    GK_HIP(rocprim::radix_sort_keys(t, tmp, in, out, (size_t)n, 0, end_bit, ctx->stream));
    return GK_OK;
}

int scan_i32(gk_ctx *ctx, Bufs &B, const int32_t *in, int32_t *out, int64_t n) {
    size_t tmp = 0;
    GK_HIP(rocprim::exclusive_scan(nullptr, tmp, in, out, 0, (size_t)n, rocprim::plus<int32_t>(), ctx->stream));
    void *t;
    GK_TRY(B.alloc((char **)&t, (int64_t)tmp));
    GK_HIP(rocprim::exclusive_scan(t, tmp, in, out, 0, (size_t)n, rocprim::plus<int32_t>(), ctx->stream));
    return GK_OK;
}

gk_dcsr *new_dcsr(gk_ctx *ctx, int64_t m, int64_t nz, int32_t *rst, int32_t *col, bool lower, bool symmetric) {
    auto *d = new gk_dcsr();
    d->ctx = ctx;
    d->m = d->n = m;
    d->nz = nz;
    d->rst = rst;
    d->col = col;
    d->owns = true;
    d->sorted = true;
    d->lower = lower;
    d->symmetric = symmetric;
    return d;
}

// Steps 2-5 on the ne generated keys in `key` (consumed).
int clean_and_build(gk_ctx *ctx, Bufs &B, int64_t n, int64_t ne, uint64_t *key, int relabel, gk_dcsr **A,
                    gk_dcsr **L, int32_t *d_perm) {
    if (n <= 0 || n > INT32_MAX - 1) {
        set_error("vertex count %lld out of int32 range", (long long)n);
        return GK_ERANGE;
    }
    uint64_t *sorted, *uniq;
    int64_t *nu;
    GK_TRY(B.alloc(&sorted, ne));
    GK_TRY(sort_keys(ctx, B, key, sorted, ne, 64));
    GK_TRY(B.alloc(&uniq, ne));
    GK_TRY(B.alloc(&nu, 1));
    {
        size_t tmp = 0;
        GK_HIP(rocprim::unique(nullptr, tmp, sorted, uniq, nu, (size_t)ne, rocprim::equal_to<uint64_t>(),
                               ctx->stream));
        void *t;
        GK_TRY(B.alloc((char **)&t, (int64_t)tmp));
        GK_HIP(rocprim::unique(t, tmp, sorted, uniq, nu, (size_t)ne, rocprim::equal_to<uint64_t>(), ctx->stream));
    }
    int64_t hn = 0;
    GK_HIP(d2h(ctx, &hn, nu, sizeof hn));
    uint64_t last = kNoEdge;
    if (hn > 0) GK_HIP(d2h(ctx, &last, uniq + hn - 1, sizeof last));
    const int64_t E = hn - (hn > 0 && last == kNoEdge ? 1 : 0);
    if (2 * E > INT32_MAX) {
        set_error("nnz %lld exceeds int32 CSR", (long long)(2 * E));
        return GK_ERANGE;
    }
    int32_t *deg, *perm, *ndeg;
    uint64_t *rk, *rks;
    GK_TRY(B.alloc(&deg, n));
    GK_HIP(hipMemsetAsync(deg, 0, sizeof(int32_t) * (size_t)n, ctx->stream));
    if (E) k_degrees<<<grid_for(ctx, E), 256, 0, ctx->stream>>>(E, uniq, deg);
    GK_TRY(B.alloc(&perm, n));
    if (relabel) {
        GK_TRY(B.alloc(&rk, n));
        GK_TRY(B.alloc(&rks, n));
        k_relabel_keys<<<grid_for(ctx, n), 256, 0, ctx->stream>>>(n, deg, rk);
        GK_TRY(sort_keys(ctx, B, rk, rks, n, 63));
        k_perm<<<grid_for(ctx, n), 256, 0, ctx->stream>>>(n, rks, 1, perm);
    } else {
        k_perm<<<grid_for(ctx, n), 256, 0, ctx->stream>>>(n, nullptr, 0, perm);
    }
    GK_TRY(B.alloc(&ndeg, n + 1));
    k_new_degrees<<<grid_for(ctx, n), 256, 0, ctx->stream>>>(n, deg, perm, ndeg);
    GK_TRY(launch_check("graph build: labels"));
    // outputs (owned by the returned handles)
    int32_t *arst = nullptr, *acol = nullptr, *lrst = nullptr, *lcol = nullptr;
    auto fail = [&](int rc) {
        (void)hipFree(arst);
        (void)hipFree(acol);
        (void)hipFree(lrst);
        (void)hipFree(lcol);
        return rc;
    };
    if (hipMalloc(&arst, 4 * (size_t)(n + 1)) != hipSuccess || hipMalloc(&acol, 4 * (size_t)std::max<int64_t>(2 * E, 1)) != hipSuccess ||
        hipMalloc(&lrst, 4 * (size_t)(n + 1)) != hipSuccess || hipMalloc(&lcol, 4 * (size_t)std::max<int64_t>(E, 1)) != hipSuccess) {
        set_error("hipMalloc of the built graph failed");
        return fail(GK_ENOMEM);
    }
    int rc = scan_i32(ctx, B, ndeg, arst, n + 1);
    if (rc != GK_OK) return fail(rc);
    if (E) {
        uint64_t *sk, *sks;
        if ((rc = B.alloc(&sk, 2 * E)) != GK_OK || (rc = B.alloc(&sks, 2 * E)) != GK_OK) return fail(rc);
        k_sym_keys<<<grid_for(ctx, E), 256, 0, ctx->stream>>>(E, uniq, perm, sk);
        if ((rc = sort_keys(ctx, B, sk, sks, 2 * E, 64)) != GK_OK) return fail(rc);
        k_low32<<<grid_for(ctx, 2 * E), 256, 0, ctx->stream>>>(2 * E, sks, acol);
    }
    int32_t *llen;
    if ((rc = B.alloc(&llen, n + 1)) != GK_OK) return fail(rc);
    k_lower_count<<<grid_for(ctx, n), 256, 0, ctx->stream>>>(n, arst, acol, llen);
    if ((rc = scan_i32(ctx, B, llen, lrst, n + 1)) != GK_OK) return fail(rc);
    k_lower_fill<<<(int)std::max<int64_t>(1, std::min<int64_t>((n + 3) / 4, (int64_t)ctx->num_cus * 16)), 256, 0,
                   ctx->stream>>>(n, arst, acol, lrst, lcol);
    if ((rc = launch_check("graph build")) != GK_OK) return fail(rc);
    if (d_perm && hipMemcpyAsync(d_perm, perm, 4 * (size_t)n, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess) {
        set_error("copy of the relabelling failed");
        return fail(GK_EHIP);
    }
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
        set_error("graph build failed on the device");
        return fail(GK_EHIP);
    }
    *A = new_dcsr(ctx, n, 2 * E, arst, acol, false, true);
    *L = new_dcsr(ctx, n, E, lrst, lcol, true, false);
    return GK_OK;
}

}  // namespace
}  // namespace gk

extern "C" int gk_dgraph_rmat(gk_ctx *ctx, int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                              int relabel, gk_dcsr **A, gk_dcsr **L, int32_t *d_perm) {
    GK_CHECK_ARG(ctx && A && L, "null argument");
    GK_CHECK_ARG(scale >= 1 && scale <= 30, "R-MAT scale %d out of range [1,30]", scale);
    GK_CHECK_ARG(edge_factor >= 1, "edge factor must be positive");
    GK_CHECK_ARG(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0, "R-MAT probabilities invalid");
    const int64_t n = int64_t(1) << scale, ne = n * edge_factor;
    Bufs B;
    uint64_t *key;
    GK_TRY(B.alloc(&key, ne));
    k_gen_rmat<<<grid_for(ctx, ne), 256, 0, ctx->stream>>>(ne, scale, a, a + b, a + b + c, seed, key);
    GK_TRY(launch_check("k_gen_rmat"));
    return clean_and_build(ctx, B, n, ne, key, relabel, A, L, d_perm);
}

extern "C" int gk_dgraph_er(gk_ctx *ctx, int64_t n, double mean_degree, uint64_t seed, int relabel, gk_dcsr **A,
                            gk_dcsr **L, int32_t *d_perm) {
    GK_CHECK_ARG(ctx && A && L, "null argument");
    GK_CHECK_ARG(n >= 2 && mean_degree > 0, "ER needs n >= 2 and a positive mean degree");
    const int64_t ne = (int64_t)std::llround(n * mean_degree / 2.0);
    Bufs B;
    uint64_t *key;
    GK_TRY(B.alloc(&key, ne));
    k_gen_er<<<grid_for(ctx, ne), 256, 0, ctx->stream>>>(ne, n, seed, key);
    GK_TRY(launch_check("k_gen_er"));
    return clean_and_build(ctx, B, n, ne, key, relabel, A, L, d_perm);
}

extern "C" int gk_dgraph_chung_lu(gk_ctx *ctx, int64_t n, double exponent, double mean_degree, uint64_t seed,
                                  int relabel, gk_dcsr **A, gk_dcsr **L, int32_t *d_perm) {
    GK_CHECK_ARG(ctx && A && L, "null argument");
    GK_CHECK_ARG(n >= 2 && exponent > 1.0 && mean_degree > 0, "Chung-Lu needs n >= 2, exponent > 1, degree > 0");
    std::vector<double> cdf;
    chung_lu_cdf(n, exponent, cdf);
    const int64_t ne = (int64_t)std::llround(n * mean_degree / 2.0);
    Bufs B;
    double *dcdf;
    uint64_t *key;
    GK_TRY(B.alloc(&dcdf, n));
    GK_HIP(h2d_chunked(ctx, dcdf, cdf.data(), sizeof(double) * (size_t)n));
    GK_TRY(B.alloc(&key, ne));
    k_gen_chung_lu<<<grid_for(ctx, ne), 256, 0, ctx->stream>>>(ne, n, dcdf, seed, key);
    GK_TRY(launch_check("k_gen_chung_lu"));
    return clean_and_build(ctx, B, n, ne, key, relabel, A, L, d_perm);
}

extern "C" int gk_dgraph_grid(gk_ctx *ctx, int64_t rows, int64_t cols, int diagonal, int relabel, gk_dcsr **A,
                              gk_dcsr **L, int32_t *d_perm) {
    GK_CHECK_ARG(ctx && A && L, "null argument");
    GK_CHECK_ARG(rows >= 1 && cols >= 1, "grid needs positive dimensions");
    const int per = diagonal ? 3 : 2;
    const int64_t n = rows * cols, ne = n * per;
    Bufs B;
    uint64_t *key;
    GK_TRY(B.alloc(&key, ne));
    k_gen_grid<<<grid_for(ctx, ne), 256, 0, ctx->stream>>>(rows, cols, per, key);
    GK_TRY(launch_check("k_gen_grid"));
    return clean_and_build(ctx, B, n, ne, key, relabel, A, L, d_perm);
}

extern "C" int gk_dcsr_download(gk_ctx *ctx, const gk_dcsr *d, int32_t *rst, int32_t *col) {
    GK_CHECK_ARG(ctx && d && rst && (col || d->nz == 0), "null argument");
    GK_HIP(d2h(ctx, rst, d->rst, sizeof(int32_t) * (size_t)(d->m + 1)));
    if (d->nz) GK_HIP(d2h(ctx, col, d->col, sizeof(int32_t) * (size_t)d->nz));
    return GK_OK;
}

extern "C" int gk_dcsr_dims(const gk_dcsr *d, int64_t *m, int64_t *n, int64_t *nz) {
    GK_CHECK_ARG(d && m && n && nz, "null argument");
    *m = d->m;
    *n = d->n;
    *nz = d->nz;
    return GK_OK;
}
