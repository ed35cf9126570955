// k-truss and full truss decomposition on gfx950.
//
// GraphKokkos's k_truss (src/main.cpp:93-141) repeats
//     Tmp<C> = C*C;  keep entries with Tmp >= k-2 (values reset to 1)
// until an iteration removes nothing, with a serial host compaction.  Here
// the graph lives in HBM as its strictly-lower part L (every undirected
// edge once, rows sorted), and one round is
//     1. supports: the masked product C<L> = L*L in SUPPORT mode, which adds
//        every triangle (j<k<i) to its three edges -- the values of
//        S<A> = A*A at A's pattern, computed on the oriented graph;
//     2. threshold: k = max(k, min support + 3) (decomposition) or the
//        fixed k (single k-truss);
//     3. prune + compact on the device: per-row keep counts, exclusive scan,
//        order-preserving scatter; removed edges get truss number k-1.
// Rounds repeat until no edge is left (decomposition) or none is removed
// (single k).  Truss numbers come back per entry of the symmetric pattern.
#include <cstring>

#include <rocprim/device/device_scan.hpp>

#include "gk_device.h"
#include "spgemm.h"

using namespace gk;

namespace gk {

int exclusive_scan_i32(gk_ctx *ctx, const int32_t *in, int32_t *out, int64_t n) {
    size_t tmp = 0;
    GK_HIP(rocprim::exclusive_scan(nullptr, tmp, in, out, 0, (size_t)n, rocprim::plus<int32_t>(), ctx->stream));
    void *t;
    GK_TRY(scratch(ctx, SC_CUB, tmp, &t));
    GK_HIP(rocprim::exclusive_scan(t, tmp, in, out, 0, (size_t)n, rocprim::plus<int32_t>(), ctx->stream));
    return GK_OK;
}

namespace {

// Length of the strictly-lower prefix of every row of a sorted symmetric
// pattern (extract_lower_triangle, include/matrix.h:698-726).
__global__ void k_lower_len(int64_t m, const int32_t *__restrict__ rst, const int32_t *__restrict__ col,
                            int32_t *__restrict__ len) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        int32_t l = rst[i], h = rst[i + 1];
        while (l < h) {
            int32_t mid = (l + h) >> 1;
            if (col[mid] < (int32_t)i) l = mid + 1; else h = mid;
        }
        len[i] = l - rst[i];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) len[m] = 0;
}

__global__ void k_lower_copy(int64_t m, const int32_t *__restrict__ rst, const int32_t *__restrict__ col,
                             const int32_t *__restrict__ lrst, int32_t *__restrict__ lcol, int32_t *__restrict__ lid) {
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < m; i += waves) {
        int32_t n = lrst[i + 1] - lrst[i];
        for (int32_t x = lane; x < n; x += 64) {
            lcol[lrst[i] + x] = col[rst[i] + x];
            lid[lrst[i] + x] = lrst[i] + x;
        }
    }
}

// Minimum support over the current edges.
__global__ void k_min_u32(int64_t n, const uint32_t *__restrict__ v, uint32_t *__restrict__ out) {
    uint32_t mn = 0xffffffffu;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        mn = min(mn, v[i]);
    mn = dev::wave_max(~mn);
    if ((threadIdx.x & 63) == 0) atomicMax(out, mn);  // out holds ~min
}

__global__ void k_keep_count(int64_t m, const int32_t *__restrict__ rst, const uint32_t *__restrict__ sup,
                             uint32_t thresh, int32_t *__restrict__ cnt) {
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < m; i += waves) {
        int32_t c = 0;
        for (int32_t p = rst[i] + lane; p < rst[i + 1]; p += 64) c += sup[p] >= thresh;
        c = dev::wave_sum(c);
        if (lane == 0) cnt[i] = c;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) cnt[m] = 0;
}

// Order-preserving compaction of one round; removed edges get `label`.
__global__ void k_keep_scatter(int64_t m, const int32_t *__restrict__ rst, const int32_t *__restrict__ col,
                               const int32_t *__restrict__ id, const uint32_t *__restrict__ sup, uint32_t thresh,
                               const int32_t *__restrict__ nrst, int32_t *__restrict__ ncol,
                               int32_t *__restrict__ nid, uint32_t *__restrict__ truss, uint32_t label) {
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < m; i += waves) {
        int32_t out = nrst[i];
        for (int32_t base = rst[i]; base < rst[i + 1]; base += 64) {
            int32_t p = base + lane;
            bool valid = p < rst[i + 1];
            bool keep = valid && sup[p] >= thresh;
            uint64_t bal = __ballot(keep);
            int32_t before = __popcll(bal & ((1ull << lane) - 1));
            if (keep) {
                ncol[out + before] = col[p];
                nid[out + before] = id[p];
            } else if (valid && truss) {
                truss[id[p]] = label;
            }
            out += __popcll(bal);
        }
    }
}

// truss number per entry of the symmetric pattern from truss per L edge.
__global__ void k_truss_to_sym(int64_t m, const int32_t *__restrict__ rst, const int32_t *__restrict__ col,
                               const int32_t *__restrict__ lrst, const uint32_t *__restrict__ tl,
                               uint32_t *__restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < m; i += waves) {
        for (int32_t p = rst[i] + lane; p < rst[i + 1]; p += 64) {
            int32_t j = col[p];
            int64_t e;
            if (j < i) {
                e = lrst[i] + (p - rst[i]);
            } else {
                // position of i in row j; L_j is the prefix of row j.
                int32_t l = rst[j], h = rst[j + 1];
                while (l < h) {
                    int32_t mid = (l + h) >> 1;
                    if (col[mid] < (int32_t)i) l = mid + 1; else h = mid;
                }
                e = lrst[j] + (l - rst[j]);
            }
            out[p] = tl[e];
        }
    }
}

int waves_grid(gk_ctx *ctx, int64_t rows) {
    int64_t g = (rows + 3) / 4;
    int64_t cap = (int64_t)ctx->num_cus * 16;
    return (int)std::max<int64_t>(1, std::min(g, cap));
}

struct Lgraph {
    int64_t m = 0, nz = 0;
    int32_t *rst = nullptr, *col = nullptr, *id = nullptr;
};

// Device buffers of a truss run: two ping-pong L graphs plus supports.
struct TrussBufs {
    gk_ctx *ctx;
    std::vector<void *> allocs;
    explicit TrussBufs(gk_ctx *c) : ctx(c) {}
    ~TrussBufs() {
        for (void *p : allocs) (void)hipFree(p);
    }
    template <class T>
    int alloc(T **p, size_t n) {
        if (hipMalloc((void **)p, sizeof(T) * (n ? n : 1)) != hipSuccess) {
            set_error("hipMalloc(%zu) in truss run failed", sizeof(T) * n);
            return GK_ENOMEM;
        }
        allocs.push_back(*p);
        return GK_OK;
    }
};

// L0 (strict lower of the symmetric A) on the device.
int build_lower(gk_ctx *ctx, const gk_dcsr *A, TrussBufs &B, Lgraph *L) {
    const int64_t m = A->m;
    int32_t *len;
    GK_TRY(B.alloc(&len, m + 1));
    GK_TRY(B.alloc(&L->rst, m + 1));
    int g = (int)std::max<int64_t>(1, std::min<int64_t>((m + 255) / 256, ctx->num_cus * 16));
    k_lower_len<<<g, 256, 0, ctx->stream>>>(m, A->rst, A->col, len);
    GK_TRY(launch_check("k_lower_len"));
    GK_TRY(exclusive_scan_i32(ctx, len, L->rst, m + 1));
    int32_t nz;
    GK_HIP(hipMemcpyAsync(&nz, L->rst + m, 4, hipMemcpyDeviceToHost, ctx->stream));
    GK_HIP(hipStreamSynchronize(ctx->stream));
    L->m = m;
    L->nz = nz;
    GK_TRY(B.alloc(&L->col, nz));
    GK_TRY(B.alloc(&L->id, nz));
    k_lower_copy<<<waves_grid(ctx, m), 256, 0, ctx->stream>>>(m, A->rst, A->col, L->rst, L->col, L->id);
    return launch_check("k_lower_copy");
}

// The round loop.  fixed_k > 0: single k-truss (stop when nothing is
// removed); fixed_k == 0: decomposition (stop when no edge is left).
// truss_l (per L0 edge) gets labels; alive graph left in *cur.
int truss_rounds(gk_ctx *ctx, TrussBufs &B, Lgraph &cur, Lgraph &nxt, int fixed_k, uint32_t *truss_l,
                 const gk_opts &o, int64_t *rounds, int32_t *kmax, double *support_ms) {
    uint32_t *sup, *dmin;
    int32_t *cnt;
    GK_TRY(B.alloc(&sup, cur.nz));
    GK_TRY(B.alloc(&cnt, cur.m + 1));
    GK_TRY(B.alloc(&dmin, 1));
    uint32_t *h;
    GK_TRY(pinned(ctx, 16, (void **)&h));
    int k = fixed_k > 0 ? fixed_k : 3;
    int32_t kbest = cur.nz ? 2 : 0;
    int64_t r = 0;
    double sms = 0;
    while (cur.nz > 0) {
        ++r;
        gk_dcsr L = view(ctx, cur.m, cur.m, cur.nz, cur.rst, cur.col, nullptr);
        ProductInfo pi;
        GK_TRY(masked_product(ctx, &L, &L, &L, MODE_SUPPORT, (int32_t *)sup, nullptr, nullptr, o, &pi));
        record_timing(ctx, pi, false);
        sms += pi.kernel_ms;
        uint32_t thresh;
        if (fixed_k > 0) {
            thresh = (uint32_t)std::max(0, fixed_k - 2);
        } else {
            GK_HIP(hipMemsetAsync(dmin, 0, 4, ctx->stream));
            int g = (int)std::max<int64_t>(1, std::min<int64_t>((cur.nz + 255) / 256, ctx->num_cus * 8));
            k_min_u32<<<g, 256, 0, ctx->stream>>>(cur.nz, sup, dmin);
            GK_TRY(launch_check("k_min_u32"));
            GK_HIP(hipMemcpyAsync(h, dmin, 4, hipMemcpyDeviceToHost, ctx->stream));
            GK_HIP(hipStreamSynchronize(ctx->stream));
            uint32_t minsup = ~h[0];
            k = std::max<int64_t>(k, (int64_t)minsup + 3);
            thresh = (uint32_t)(k - 2);
        }
        const int g = waves_grid(ctx, cur.m);
        k_keep_count<<<g, 256, 0, ctx->stream>>>(cur.m, cur.rst, sup, thresh, cnt);
        GK_TRY(launch_check("k_keep_count"));
        if (!nxt.rst) GK_TRY(B.alloc(&nxt.rst, cur.m + 1));
        GK_TRY(exclusive_scan_i32(ctx, cnt, nxt.rst, cur.m + 1));
        int32_t nz;
        GK_HIP(hipMemcpyAsync(&h[1], nxt.rst + cur.m, 4, hipMemcpyDeviceToHost, ctx->stream));
        GK_HIP(hipStreamSynchronize(ctx->stream));
        nz = (int32_t)h[1];
        if (nz == cur.nz) {
            // nothing removed: only possible for a fixed k (the
            // decomposition threshold always removes the minimum).
            break;
        }
        if (!nxt.col) {
            GK_TRY(B.alloc(&nxt.col, cur.nz));
            GK_TRY(B.alloc(&nxt.id, cur.nz));
        }
        k_keep_scatter<<<g, 256, 0, ctx->stream>>>(cur.m, cur.rst, cur.col, cur.id, sup, thresh, nxt.rst, nxt.col,
                                                   nxt.id, truss_l, (uint32_t)(k - 1));
        GK_TRY(launch_check("k_keep_scatter"));
        if (fixed_k == 0) kbest = k - 1;
        nxt.m = cur.m;
        nxt.nz = nz;
        std::swap(cur, nxt);
    }
    if (rounds) *rounds = r;
    if (kmax) *kmax = kbest;
    if (support_ms) *support_ms = sms;
    return GK_OK;
}

}  // namespace
}  // namespace gk

extern "C" int gk_truss_decompose_dev(gk_ctx *ctx, const gk_dcsr *A, uint32_t *d_truss, int32_t *kmax,
                                      const gk_opts *opts) {
    GK_CHECK_ARG(ctx && A && (d_truss || A->nz == 0), "null argument");
    GK_CHECK_ARG(A->m == A->n, "A must be square");
    gk_opts o = resolve_opts(opts);
    ctx->last = gk_timing{};
    Timer total(ctx);
    total.start();
    TrussBufs B(ctx);
    Lgraph cur, nxt;
    GK_TRY(build_lower(ctx, A, B, &cur));
    const int64_t nl = cur.nz;
    int32_t *lrst0 = cur.rst;
    // Keep L0's row pointers: the final mapping to symmetric entries needs
    // them, while `cur` is replaced every round.
    int32_t *lrst_keep;
    GK_TRY(B.alloc(&lrst_keep, A->m + 1));
    GK_HIP(hipMemcpyAsync(lrst_keep, lrst0, 4 * (A->m + 1), hipMemcpyDeviceToDevice, ctx->stream));
    uint32_t *tl;
    GK_TRY(B.alloc(&tl, nl));
    int64_t rounds = 0;
    int32_t kb = 0;
    double sms = 0;
    GK_TRY(truss_rounds(ctx, B, cur, nxt, 0, tl, o, &rounds, &kb, &sms));
    if (A->nz) {
        k_truss_to_sym<<<waves_grid(ctx, A->m), 256, 0, ctx->stream>>>(A->m, A->rst, A->col, lrst_keep, tl, d_truss);
        GK_TRY(launch_check("k_truss_to_sym"));
    }
    total.stop();
    ctx->last.total_ms = total.ms();
    ctx->last.iterations = rounds;
    if (kmax) *kmax = kb;
    return GK_OK;
}

extern "C" int gk_truss_decompose(gk_ctx *ctx, const gk_csr *A, uint32_t *truss, int32_t *kmax, const gk_opts *opts) {
    GK_CHECK_ARG(ctx && A && (truss || A->nz == 0), "null argument");
    gk_dcsr d;
    GK_TRY(upload_csr(ctx, A, &d));
    uint32_t *dt = nullptr;
    if (hipMalloc(&dt, 4 * (A->nz ? A->nz : 1)) != hipSuccess) { free_csr(&d); set_error("hipMalloc failed"); return GK_ENOMEM; }
    int rc = gk_truss_decompose_dev(ctx, &d, dt, kmax, opts);
    if (rc == GK_OK && A->nz && d2h(ctx, truss, dt, 4 * A->nz) != hipSuccess) {
        set_error("copy of truss numbers failed");
        rc = GK_EHIP;
    }
    (void)hipFree(dt);
    free_csr(&d);
    return rc;
}

namespace gk {
namespace {
// keep[p] per symmetric entry from the alive L edges (tl[e] == 0: alive).
__global__ void k_alive_to_sym(int64_t m, const int32_t *__restrict__ rst, const int32_t *__restrict__ col,
                               const int32_t *__restrict__ lrst, const uint32_t *__restrict__ tl,
                               int32_t *__restrict__ keep_cnt, uint8_t *__restrict__ keep) {
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < m; i += waves) {
        int32_t c = 0;
        for (int32_t p = rst[i] + lane; p < rst[i + 1]; p += 64) {
            int32_t j = col[p];
            int64_t e;
            if (j < i) {
                e = lrst[i] + (p - rst[i]);
            } else {
                int32_t l = rst[j], h = rst[j + 1];
                while (l < h) {
                    int32_t mid = (l + h) >> 1;
                    if (col[mid] < (int32_t)i) l = mid + 1; else h = mid;
                }
                e = lrst[j] + (l - rst[j]);
            }
            uint8_t k = tl[e] == 0;
            keep[p] = k;
            c += k;
        }
        c = dev::wave_sum(c);
        if (lane == 0) keep_cnt[i] = c;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) keep_cnt[m] = 0;
}

__global__ void k_keep_sym_scatter(int64_t m, const int32_t *__restrict__ rst, const int32_t *__restrict__ col,
                                   const uint8_t *__restrict__ keep, const int32_t *__restrict__ nrst,
                                   int32_t *__restrict__ ncol) {
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < m; i += waves) {
        int32_t out = nrst[i];
        for (int32_t base = rst[i]; base < rst[i + 1]; base += 64) {
            int32_t p = base + lane;
            bool k = p < rst[i + 1] && keep[p];
            uint64_t bal = __ballot(k);
            if (k) ncol[out + __popcll(bal & ((1ull << lane) - 1))] = col[p];
            out += __popcll(bal);
        }
    }
}
}  // namespace
}  // namespace gk

extern "C" int gk_ktruss(gk_ctx *ctx, const gk_csr *A, int k, int32_t *rst_out, int32_t *col_out, int64_t *nz_out,
                         int64_t *iterations, const gk_opts *opts) {
    GK_CHECK_ARG(ctx && A && rst_out && nz_out, "null argument");
    GK_CHECK_ARG(A->m == A->n, "A must be square");
    GK_CHECK_ARG(k >= 2, "k must be at least 2");
    gk_opts o = resolve_opts(opts);
    gk_dcsr d;
    GK_TRY(upload_csr(ctx, A, &d));
    int rc = GK_OK;
    {
        TrussBufs B(ctx);
        Lgraph cur, nxt;
        int64_t rounds = 0;
        uint32_t *tl = nullptr;
        int32_t *lrst_keep = nullptr, *cnt = nullptr, *nrst = nullptr, *ncol = nullptr;
        uint8_t *keep = nullptr;
        do {
            if ((rc = build_lower(ctx, &d, B, &cur))) break;
            if ((rc = B.alloc(&lrst_keep, A->m + 1))) break;
            if (hipMemcpyAsync(lrst_keep, cur.rst, 4 * (A->m + 1), hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess) {
                rc = GK_EHIP; set_error("copy failed"); break;
            }
            if ((rc = B.alloc(&tl, cur.nz))) break;
            if ((rc = zero_u32(ctx, tl, cur.nz))) break;
            Timer t(ctx);
            ctx->last = gk_timing{};
            t.start();
            // Removed edges are labelled 1; alive ones stay 0.
            if ((rc = truss_rounds(ctx, B, cur, nxt, k, tl, o, &rounds, nullptr, nullptr))) break;
            t.stop();
            ctx->last.total_ms = t.ms();
            ctx->last.iterations = rounds;
            if ((rc = B.alloc(&cnt, A->m + 1)) || (rc = B.alloc(&nrst, A->m + 1)) || (rc = B.alloc(&keep, A->nz)) ||
                (rc = B.alloc(&ncol, A->nz)))
                break;
            k_alive_to_sym<<<waves_grid(ctx, A->m), 256, 0, ctx->stream>>>(A->m, d.rst, d.col, lrst_keep, tl, cnt, keep);
            if ((rc = launch_check("k_alive_to_sym"))) break;
            if ((rc = exclusive_scan_i32(ctx, cnt, nrst, A->m + 1))) break;
            k_keep_sym_scatter<<<waves_grid(ctx, A->m), 256, 0, ctx->stream>>>(A->m, d.rst, d.col, keep, nrst, ncol);
            if ((rc = launch_check("k_keep_sym_scatter"))) break;
            if (d2h(ctx, rst_out, nrst, 4 * (A->m + 1)) != hipSuccess) { rc = GK_EHIP; set_error("copy failed"); break; }
            *nz_out = rst_out[A->m];
            if (col_out && *nz_out &&
                d2h(ctx, col_out, ncol, 4 * (size_t)*nz_out) != hipSuccess) {
                rc = GK_EHIP; set_error("copy failed"); break;
            }
            if (iterations) *iterations = rounds;
        } while (0);
    }
    free_csr(&d);
    return rc;
}

extern "C" int gk_ktruss_support(gk_ctx *ctx, const gk_csr *A, uint32_t *support, const gk_opts *opts) {
    GK_CHECK_ARG(ctx && A && (support || A->nz == 0), "null argument");
    GK_CHECK_ARG(A->m == A->n, "A must be square");
    gk_opts o = resolve_opts(opts);
    gk_dcsr d;
    GK_TRY(upload_csr(ctx, A, &d));
    int rc = GK_OK;
    {
        TrussBufs B(ctx);
        Lgraph L;
        uint32_t *sup = nullptr, *dout = nullptr;
        int32_t *lrst_keep = nullptr;
        do {
            if ((rc = build_lower(ctx, &d, B, &L))) break;
            if ((rc = B.alloc(&sup, L.nz)) || (rc = B.alloc(&dout, A->nz))) break;
            gk_dcsr Ld = view(ctx, L.m, L.m, L.nz, L.rst, L.col, nullptr);
            ProductInfo pi;
            if ((rc = masked_product(ctx, &Ld, &Ld, &Ld, MODE_SUPPORT, (int32_t *)sup, nullptr, nullptr, o, &pi))) break;
            lrst_keep = L.rst;
            if (A->nz) {
                k_truss_to_sym<<<waves_grid(ctx, A->m), 256, 0, ctx->stream>>>(A->m, d.rst, d.col, lrst_keep, sup, dout);
                if ((rc = launch_check("k_truss_to_sym"))) break;
                if (d2h(ctx, support, dout, 4 * A->nz) != hipSuccess) { rc = GK_EHIP; set_error("copy failed"); break; }
            }
            record_timing(ctx, pi, true);
        } while (0);
    }
    free_csr(&d);
    return rc;
}
