// Context, device CSR handles, scratch buffers and timing for libgraphhip.so.
// The context is this library's counterpart of Kokkos::initialize
// (src/main.cpp:358 in GraphKokkos): it binds one device and one stream.
#include <cstring>

#include "gk_device.h"

using namespace gk;

namespace gk {

int scratch(gk_ctx *ctx, int s, size_t bytes, void **out) {
    if (bytes == 0) bytes = 256;
    if (ctx->scratch_bytes[s] < bytes) {
        if (ctx->scratch[s]) {
            GK_HIP(hipStreamSynchronize(ctx->stream));
            GK_HIP(hipFree(ctx->scratch[s]));
            ctx->scratch[s] = nullptr;
            ctx->scratch_bytes[s] = 0;
        }
        size_t want = bytes + bytes / 4;  // headroom against regrowth
        hipError_t e = hipMalloc(&ctx->scratch[s], want);
        if (e != hipSuccess) {
            set_error("hipMalloc(%zu) for scratch slot %d failed: %s", want, s, hipGetErrorString(e));
            return GK_ENOMEM;
        }
        ctx->scratch_bytes[s] = want;
    }
    *out = ctx->scratch[s];
    return GK_OK;
}

int pinned(gk_ctx *ctx, size_t bytes, void **out) {
    static const size_t kPinned = 1 << 16;
    if (bytes > kPinned) { set_error("pinned staging request too large"); return GK_EINVAL; }
    if (!ctx->pinned) GK_HIP(hipHostMalloc(&ctx->pinned, kPinned, hipHostMallocDefault));
    *out = ctx->pinned;
    return GK_OK;
}

gk_dcsr view(gk_ctx *ctx, int64_t m, int64_t n, int64_t nz, int32_t *rst, int32_t *col, int32_t *val) {
    gk_dcsr d;
    d.ctx = ctx; d.m = m; d.n = n; d.nz = nz; d.rst = rst; d.col = col; d.val = val; d.owns = false;
    return d;
}

// Device-to-host copy on the context stream, in pieces of at most 256 MiB
// (see h2d_chunked).
hipError_t d2h(gk_ctx *ctx, void *dst, const void *src, size_t bytes) {
    constexpr size_t kChunk = size_t(256) << 20;
    for (size_t off = 0; off < bytes; off += kChunk) {
        const size_t n = std::min(kChunk, bytes - off);
        hipError_t e = hipMemcpyAsync((char *)dst + off, (const char *)src + off, n, hipMemcpyDeviceToHost,
                                      ctx->stream);
        if (e != hipSuccess) return e;
        e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

int check_csr(const gk_csr *c, const char *name) {
    if (!c) { set_error("%s is NULL", name); return GK_EINVAL; }
    if (c->m < 0 || c->n < 0 || c->nz < 0) { set_error("%s has negative sizes", name); return GK_EINVAL; }
    if (c->m >= INT32_MAX || c->n >= INT32_MAX || c->nz >= INT32_MAX) {
        set_error("%s exceeds int32 CSR indexing", name);
        return GK_ERANGE;
    }
    if (!c->rst || (c->nz > 0 && !c->col)) { set_error("%s has NULL arrays", name); return GK_EINVAL; }
    if (c->rst[0] != 0 || c->rst[c->m] != c->nz) {
        set_error("%s: Rst[0] must be 0 and Rst[m] must equal nz", name);
        return GK_EINVAL;
    }
    return GK_OK;
}

int inspect_csr(const gk_csr *c, const char *name, gk_dcsr *d) {
    GK_TRY(check_csr(c, name));
    const int64_t m = c->m, n = c->n;
    const int32_t *rst = c->rst, *col = c->col;
    std::atomic<int64_t> bad_row{-1}, bad_col{-1};
    std::atomic<bool> unsorted{false}, upper{false};
    // pass 1: row pointers, column range, row order, triangle side
    par_dynamic(m, 4096, [&](int64_t i) {
        const int32_t b = rst[i], e = rst[i + 1];
        if (e < b || b < 0 || e > c->nz) { bad_row.store(i); return; }
        int32_t prev = -1;
        bool srt = true, low = true;
        for (int32_t p = b; p < e; ++p) {
            const int32_t j = col[p];
            if (j < 0 || j >= n) { bad_col.store(p); return; }
            srt &= j > prev;
            low &= j < i;
            prev = j;
        }
        if (!srt) unsorted.store(true, std::memory_order_relaxed);
        if (!low) upper.store(true, std::memory_order_relaxed);
    });
    if (bad_row.load() >= 0) {
        set_error("%s: row pointers of row %lld are decreasing or outside [0, nz]", name, (long long)bad_row.load());
        return GK_EINVAL;
    }
    if (bad_col.load() >= 0) {
        set_error("%s: column index at entry %lld is outside [0, %lld)", name, (long long)bad_col.load(),
                  (long long)n);
        return GK_EINVAL;
    }
    d->sorted = !unsorted.load();
    d->lower = !upper.load();
    d->symmetric = false;
    if (d->sorted && m == n) {
        // pass 2: every (i, j) has its mirror (j, i), no diagonal entries
        std::atomic<bool> asym{false};
        par_dynamic(m, 4096, [&](int64_t i) {
            if (asym.load(std::memory_order_relaxed)) return;
            for (int32_t p = rst[i]; p < rst[i + 1]; ++p) {
                const int32_t j = col[p];
                if (j == i || !std::binary_search(col + rst[j], col + rst[j + 1], (int32_t)i)) {
                    asym.store(true, std::memory_order_relaxed);
                    return;
                }
            }
        });
        d->symmetric = !asym.load();
    }
    return GK_OK;
}

int require_sorted(const gk_dcsr *d, const char *name) {
    if (!d->sorted) {
        set_error("%s must have strictly increasing column indices in every row (sorted, no duplicates)", name);
        return GK_EINVAL;
    }
    return GK_OK;
}

int require_symmetric(const gk_dcsr *d, const char *name) {
    if (!d->symmetric) {
        set_error("%s must be the sorted, symmetric pattern of an undirected graph without self loops", name);
        return GK_EINVAL;
    }
    return GK_OK;
}

// Host-to-device copy of a host array in pieces of at most kCopyChunk
// bytes, each a plain pageable copy of the size the rest of the library
// issues, so multi-GB arrays take the same runtime path as small ones.
constexpr size_t kCopyChunk = size_t(256) << 20;

hipError_t h2d_chunked(gk_ctx *ctx, void *dst, const void *src, size_t bytes) {
    for (size_t off = 0; off < bytes; off += kCopyChunk) {
        const size_t n = std::min(kCopyChunk, bytes - off);
        hipError_t e = hipMemcpyAsync((char *)dst + off, (const char *)src + off, n, hipMemcpyHostToDevice,
                                      ctx->stream);
        if (e != hipSuccess) return e;
        e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

int upload_csr(gk_ctx *ctx, const gk_csr *h, gk_dcsr *d) {
    GK_TRY(inspect_csr(h, "csr", d));
    d->ctx = ctx; d->m = h->m; d->n = h->n; d->nz = h->nz; d->owns = true;
    d->rst = d->col = d->val = nullptr;
    GK_HIP(hipMalloc(&d->rst, sizeof(int32_t) * (h->m + 1)));
    GK_HIP(hipMalloc(&d->col, sizeof(int32_t) * (h->nz ? h->nz : 1)));
    GK_HIP(h2d_chunked(ctx, d->rst, h->rst, sizeof(int32_t) * (h->m + 1)));
    GK_HIP(h2d_chunked(ctx, d->col, h->col, sizeof(int32_t) * h->nz));
    if (h->val) {
        GK_HIP(hipMalloc(&d->val, sizeof(int32_t) * (h->nz ? h->nz : 1)));
        GK_HIP(h2d_chunked(ctx, d->val, h->val, sizeof(int32_t) * h->nz));
    }
    return GK_OK;
}

void free_csr(gk_dcsr *d) {
    if (!d || !d->owns) return;
    if (d->rst) (void)hipFree(d->rst);
    if (d->col) (void)hipFree(d->col);
    if (d->val) (void)hipFree(d->val);
    d->rst = d->col = d->val = nullptr;
}

Timer::Timer(gk_ctx *c, hipStream_t stream) : ctx(c), s(stream ? stream : c->stream), a(nullptr), b(nullptr) {
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
}
Timer::~Timer() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
}
void Timer::start() { (void)hipEventRecord(a, s); }
void Timer::stop() { (void)hipEventRecord(b, s); }
double Timer::ms() {
    float t = 0;
    (void)hipEventSynchronize(b);
    (void)hipEventElapsedTime(&t, a, b);
    return t;
}

gk_opts resolve_opts(const gk_opts *o) {
    gk_opts r;
    gk_opts_default(&r);
    if (o && o->struct_size >= (int32_t)sizeof(gk_opts)) r = *o;
    // values outside enum gk_selector (the engine's internal ones included)
    // select the fixed thresholds, as they always have
    if (r.selector != GK_SEL_LEARNED && r.selector != GK_SEL_FORCED) r.selector = GK_SEL_HEURISTIC;
    return r;
}

int launch_check(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("launch of %s failed: %s", what, hipGetErrorString(e));
        return GK_EHIP;
    }
    return GK_OK;
}

}  // namespace gk

extern "C" void gk_opts_default(gk_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->struct_size = sizeof(gk_opts);
    o->selector = GK_SEL_HEURISTIC;
    o->forced_accum = GK_ACC_MSA_LDS;
    o->mca_max_mask = -1;  // the product mode's default (mode_opts)
    o->lds_max_mask = 1536;
    o->dense_min_mask = 16384;
    o->split_flop = 1 << 18;
    o->block_threads = 64;
    o->small_max_mask = 32;
    o->blocks_small = 24;
    o->blocks_long = 0;    // the product mode's default
    o->blocks_mca = 0;     // the product mode's default; the batching step's 32
    o->blocks_global = 4;
    o->batch_blocks = 8;
    o->batch_order = 0;
    o->hub_window = 32768;
    o->learned_margin = 250;
    o->hub_probe = -1;     // the product mode's default
    o->selector_labels = 1;
}

namespace gk { int bc_noalloc_setup(gk_ctx *ctx); }

extern "C" int gk_init(int device, gk_ctx **out) {
    GK_CHECK_ARG(out, "out is NULL");
    int count = 0;
    GK_HIP(hipGetDeviceCount(&count));
    GK_CHECK_ARG(device >= 0 && device < count, "device %d not present (%d devices)", device, count);
    GK_HIP(hipSetDevice(device));
    auto *ctx = new gk_ctx();
    ctx->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) { delete ctx; set_error("hipGetDeviceProperties failed"); return GK_EHIP; }
    ctx->num_cus = prop.multiProcessorCount;
    bool ok = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; ok && i < gk_ctx::kAux; ++i)
        ok = hipStreamCreateWithFlags(&ctx->aux[i], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&ctx->join[i], hipEventDisableTiming) == hipSuccess;
    ok = ok && hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&ctx->side_fork, hipEventDisableTiming) == hipSuccess &&
         hipEventCreateWithFlags(&ctx->side_join, hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        delete ctx;
        set_error("stream or event creation failed");
        return GK_EHIP;
    }
    // Set-up of the allocation-free entry point: the pinned staging buffer
    // and this device's scratch array, so that its calls allocate nothing.
    void *staging = nullptr;
    int rc = pinned(ctx, 16, &staging);
    if (rc == GK_OK) rc = bc_noalloc_setup(ctx);
    if (rc != GK_OK) {
        gk_finalize(ctx);
        return rc;
    }
    *out = ctx;
    return GK_OK;
}

namespace gk { void model_free(Model *m); }

extern "C" void gk_finalize(gk_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (int s = 0; s < SC_NUM; ++s)
        if (ctx->scratch[s]) (void)hipFree(ctx->scratch[s]);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->dense_map) (void)hipFree(ctx->dense_map);
    for (auto &b : ctx->truss_cache)
        if (b.first) (void)hipFree(b.first);
    for (auto &b : ctx->cmask_cache)
        if (b.first) (void)hipFree(b.first);
    if (ctx->model) model_free(ctx->model);
    for (int i = 0; i < gk_ctx::kAux; ++i) {
        if (ctx->aux[i]) (void)hipStreamSynchronize(ctx->aux[i]);
        if (ctx->aux[i]) (void)hipStreamDestroy(ctx->aux[i]);
        if (ctx->join[i]) (void)hipEventDestroy(ctx->join[i]);
    }
    if (ctx->fork) (void)hipEventDestroy(ctx->fork);
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    if (ctx->side_fork) (void)hipEventDestroy(ctx->side_fork);
    if (ctx->side_join) (void)hipEventDestroy(ctx->side_join);
    for (hipEvent_t e : ctx->events) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" int gk_sync(gk_ctx *ctx) {
    GK_CHECK_ARG(ctx, "ctx is NULL");
    GK_HIP(hipStreamSynchronize(ctx->stream));
    return GK_OK;
}

extern "C" int gk_last_timing(gk_ctx *ctx, gk_timing *out) {
    GK_CHECK_ARG(ctx && out, "null argument");
    *out = ctx->last;
    return GK_OK;
}

extern "C" int gk_dcsr_upload(gk_ctx *ctx, const gk_csr *h, gk_dcsr **out) {
    GK_CHECK_ARG(ctx && out, "null argument");
    auto *d = new gk_dcsr();
    int rc = upload_csr(ctx, h, d);
    if (rc != GK_OK) { free_csr(d); delete d; return rc; }
    *out = d;
    return GK_OK;
}

extern "C" void gk_dcsr_free(gk_dcsr *d) {
    if (!d) return;
    free_csr(d);
    delete d;
}

extern "C" int64_t gk_dcsr_nnz(const gk_dcsr *d) { return d ? d->nz : -1; }

extern "C" int gk_dmalloc(gk_ctx *ctx, size_t bytes, void **dptr) {
    GK_CHECK_ARG(ctx && dptr, "null argument");
    GK_HIP(hipMalloc(dptr, bytes ? bytes : 1));
    return GK_OK;
}
extern "C" void gk_dfree(gk_ctx *ctx, void *dptr) {
    (void)ctx;
    if (dptr) (void)hipFree(dptr);
}
extern "C" int gk_memcpy_d2h(gk_ctx *ctx, void *dst, const void *src, size_t bytes) {
    GK_CHECK_ARG(ctx, "ctx is NULL");
    GK_HIP(d2h(ctx, dst, src, bytes));
    return GK_OK;
}
extern "C" int gk_memcpy_d2d(gk_ctx *ctx, void *dst, const void *src, size_t bytes) {
    GK_CHECK_ARG(ctx, "ctx is NULL");
    GK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    GK_HIP(hipStreamSynchronize(ctx->stream));
    return GK_OK;
}
extern "C" int gk_memcpy_h2d(gk_ctx *ctx, void *dst, const void *src, size_t bytes) {
    GK_CHECK_ARG(ctx, "ctx is NULL");
    GK_HIP(h2d_chunked(ctx, dst, src, bytes));
    return GK_OK;
}
