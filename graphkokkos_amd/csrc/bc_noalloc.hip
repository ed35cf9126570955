// gk_bc_batch_dev_noalloc: the batched Brandes sweeps of gk_bc_batch_dev
// (same kernels, same results) with every work buffer carved out of a
// fixed device array of GK_BC_NOALLOC_SCRATCH_BYTES, so a call allocates
// no device memory.  For callers that cannot allocate: code under HIP graph
// capture, serving loops, harnesses that forbid hipMalloc.
//
// The array is one file-scope __device__ symbol; every GPU the library runs
// on holds its own copy until the process ends.  Its name is not a device
// address in host code: gk_init calls bc_noalloc_setup, which looks up the
// current device's copy with hipGetSymbolAddress once per device, and every
// call passes that address on.  Calls on one GPU share its copy, so calls
// that overlap on one GPU race; each call initialises what it reads (the
// visited and first-level bitmaps, the counters, the scores).
#include <mutex>

#include "bc.h"
#include "gk_device.h"

using namespace gk;

static __device__ __align__(256) unsigned char scratch[GK_BC_NOALLOC_SCRATCH_BYTES];

namespace gk {
namespace {

constexpr size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Consecutive 256-byte-aligned pieces of the array, in the order bc::sweeps
// asks for its buffers; the BFS levels take the rest, one n x 4-word bitmap
// each.
struct ArenaPool final : bc::Pool {
    char *base;
    size_t cap, used = 0;
    int64_t n;
    char *levels = nullptr;  // level d at levels + d * level_bytes()
    ArenaPool(void *b, size_t c, int64_t n_) : base((char *)b), cap(c), n(n_) {}
    size_t level_bytes() const { return align256((size_t)std::max<int64_t>(n, 1) * 32); }
    int get(size_t bytes, void **p) override {
        if (levels) {
            set_error("no-alloc BC: a buffer was requested after the BFS levels");
            return GK_ESTATE;
        }
        const size_t b = align256(bytes);
        if (used + b > cap) {
            set_error("no-alloc BC: %zu more bytes needed beyond the %zu-byte scratch array (%zu used)", b, cap, used);
            return GK_ERANGE;
        }
        *p = base + used;
        used += b;
        return GK_OK;
    }
    int level(int d, uint64_t **p) override {
        if (!levels) levels = base + used;
        const size_t end = (size_t)(levels - base) + (size_t)(d + 1) * level_bytes();
        if (end > cap) {
            set_error("no-alloc BC: BFS level %d of a graph with %lld vertices does not fit the %zu-byte scratch "
                      "array (each level takes %zu bytes)", d, (long long)n, cap, level_bytes());
            return GK_ERANGE;
        }
        *p = (uint64_t *)(levels + (size_t)d * level_bytes());
        return GK_OK;
    }
};

// Bytes bc::sweeps takes before its levels on a sorted pattern with n
// vertices whatever its degrees: sigma / W (n x 256 FP64), visited bits
// (n x 4 words), the batch's sources, the level counter, the two
// heavy-vertex lists and their counts, the chunk counts.  The chunks of
// heavy vertices (kHeavy neighbours or more) come on top.
size_t fixed_bytes(int64_t n) {
    const size_t m = (size_t)std::max<int64_t>(n, 1);
    return align256(m * bc::kSrc * 8) + align256(m * 32) + align256(4 * bc::kSrc) + align256(8) +
           2 * align256(m * 4) + align256(8) + align256(16);
}

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

}  // namespace

// Called by gk_init with ctx->device current.
int bc_noalloc_setup(gk_ctx *ctx) {
    GK_CHECK_ARG(ctx->device >= 0 && ctx->device < kMaxDevices, "device %d beyond %d", ctx->device, kMaxDevices);
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[ctx->device]) {
        void *p = nullptr;
        GK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch)));  // the array above, not gk::scratch()
        g_scratch_ptr[ctx->device] = p;
    }
    ctx->bc_scratch = g_scratch_ptr[ctx->device];
    return GK_OK;
}

}  // namespace gk

extern "C" int gk_bc_batch_dev_noalloc(gk_ctx *ctx, const gk_dcsr *A, const gk_dcsr *AT, const int32_t *sources,
                                       int64_t nsrc, double *d_bc) {
    GK_TRY(bc::check_args(ctx, A, AT, sources, nsrc, d_bc));
    GK_CHECK_ARG(ctx->bc_scratch, "no-alloc BC: the context has no scratch array");
    GK_CHECK_ARG(ctx->pinned, "no-alloc BC: the context has no pinned staging buffer");
    GK_CHECK_ARG(A->sorted, "no-alloc BC needs A's rows sorted and free of repeats");
    GK_CHECK_ARG(AT || A->symmetric, "no-alloc BC needs AT or a symmetric A (it builds no transpose)");
    const int64_t n = A->m;
    ArenaPool P(ctx->bc_scratch, GK_BC_NOALLOC_SCRATCH_BYTES, n);
    const size_t need = fixed_bytes(n) + 2 * P.level_bytes();
    if (need > (size_t)GK_BC_NOALLOC_SCRATCH_BYTES) {
        set_error("no-alloc BC: %lld vertices need at least %zu bytes, beyond the %zu-byte scratch array",
                  (long long)n, need, (size_t)GK_BC_NOALLOC_SCRATCH_BYTES);
        return GK_ERANGE;
    }
    const int32_t *trst = AT ? AT->rst : A->rst, *tcol = AT ? AT->col : A->col;
    ctx->last = gk_timing{};
    return bc::sweeps(ctx, A, trst, tcol, sources, nsrc, d_bc, P, nullptr, false);
}
