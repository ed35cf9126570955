// Device-side internals of libgraphhip.so shared by the .hip translation
// units: the context, device CSR handles, scratch buffers, HIP error
// checks and small wave-level helpers for gfx950 (wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "gk_internal.h"

#define GK_HIP(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            gk::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                          __LINE__);                                                    \
            return GK_EHIP;                                                             \
        }                                                                               \
    } while (0)

#define GK_TRY(expr)              \
    do {                          \
        int rc_ = (expr);         \
        if (rc_ != GK_OK) return rc_; \
    } while (0)

namespace gk {

// Named grow-only device buffers owned by a context.
enum ScratchSlot {
    SC_FEAT = 0,      // per-row features
    SC_TASK,          // task list
    SC_BINS,          // bin counters
    SC_ACC,           // global accumulators (hash/dense)
    SC_RED,           // reductions
    SC_TMP0, SC_TMP1, SC_TMP2, SC_TMP3, SC_TMP4, SC_TMP5, SC_TMP6, SC_TMP7,
    SC_CUB,           // rocprim temporary storage
    SC_MODEL,         // selector network parameters
    SC_HUB,           // triangle counting: bitmaps of the hub rows of B
    SC_NUM
};

struct Model;  // learned selector (selector.hip)

}  // namespace gk

struct gk_ctx {
    int device = 0;
    int num_cus = 0;
    hipStream_t stream = nullptr;
    // Auxiliary streams (default priority) that run the accumulator kernels
    // of one product side by side; forked from and joined into `stream`.
    static constexpr int kAux = 4;
    hipStream_t aux[kAux] = {};
    hipEvent_t fork = nullptr, join[kAux] = {};
    // A side stream (default priority) for work that runs beside a product
    // on the streams above, forked from and joined into `stream`.
    hipStream_t side = nullptr;
    hipEvent_t side_fork = nullptr, side_join = nullptr;
    void *scratch[gk::SC_NUM] = {};
    size_t scratch_bytes[gk::SC_NUM] = {};
    void *pinned = nullptr;          // small pinned host staging buffer
    int32_t *dense_map = nullptr;    // dense-global maps, all zero between launches
    int64_t dense_blocks = 0, dense_ncols = 0;
    gk_timing last{};
    std::vector<hipEvent_t> events;  // reusable timing events
    gk::Model *model = nullptr;
    // Work buffers of the k-truss runs, kept between calls (grow-only, freed
    // by gk_finalize): the n-th allocation of a run reuses the n-th buffer.
    std::vector<std::pair<void *, size_t>> truss_cache;
    bool truss_cache_busy = false;
    // The same for the work buffers of the complement-mask product.
    std::vector<std::pair<void *, size_t>> cmask_cache;
    bool cmask_cache_busy = false;
    // Device address of this device's fixed scratch array of the
    // allocation-free BC entry point (bc_noalloc.hip), set by gk_init.
    void *bc_scratch = nullptr;
};

struct gk_dcsr {
    gk_ctx *ctx = nullptr;
    int64_t m = 0, n = 0, nz = 0;
    int32_t *rst = nullptr;
    int32_t *col = nullptr;
    int32_t *val = nullptr;   // nullptr: pattern (all ones)
    bool owns = true;
    // Structure recorded when the host arrays were checked at upload
    // (internal views are built sorted by construction):
    bool sorted = true;       // every row strictly increasing
    bool lower = false;       // every column below its row
    bool symmetric = false;   // square, pattern symmetric, no self loops
    // Capacity of rst (entries) and of col / val of a CSR the library built
    // for a result that later calls may refill (0: not refillable).
    int64_t cap_rst = 0, cap_nz = 0;
};

namespace gk {

// Device scratch of at least `bytes` in slot `s` (contents undefined).
int scratch(gk_ctx *ctx, int s, size_t bytes, void **out);
// Pinned host staging of at least `bytes`.
int pinned(gk_ctx *ctx, size_t bytes, void **out);
// Stream-ordered device-to-host copy on the context stream, then wait.
// (A plain hipMemcpy runs on the null stream, which does not wait for the
// context's non-blocking stream.)
hipError_t d2h(gk_ctx *ctx, void *dst, const void *src, size_t bytes);
// Non-owning device CSR view.
gk_dcsr view(gk_ctx *ctx, int64_t m, int64_t n, int64_t nz, int32_t *rst, int32_t *col, int32_t *val);
// Upload helpers used by the host-pointer entry points.
int upload_csr(gk_ctx *ctx, const gk_csr *h, gk_dcsr *d);
hipError_t h2d_chunked(gk_ctx *ctx, void *dst, const void *src, size_t bytes);
void free_csr(gk_dcsr *d);
int check_csr(const gk_csr *c, const char *name);
// check_csr plus a scan of the arrays: row pointers non-decreasing, columns
// in [0, n); fills d's sorted / lower / symmetric flags.
int inspect_csr(const gk_csr *c, const char *name, gk_dcsr *d);
// Entry-point preconditions on device handles.
int require_sorted(const gk_dcsr *d, const char *name);
int require_symmetric(const gk_dcsr *d, const char *name);
// Event timer on the context stream.
struct Timer {
    gk_ctx *ctx;
    hipStream_t s;
    hipEvent_t a, b;
    explicit Timer(gk_ctx *c, hipStream_t stream = nullptr);
    ~Timer();
    void start();
    void stop();
    double ms();  // synchronises on the stop event
};
// Resolve options (NULL -> defaults).
gk_opts resolve_opts(const gk_opts *o);
// Kernel-launch error check.
int launch_check(const char *what);

}  // namespace gk

// ---- device helpers -------------------------------------------------------
namespace gk {
namespace dev {

__device__ __forceinline__ int lane_id() { return __lane_id(); }

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <class T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        T w = __shfl_xor(v, o, 64);
        v = v > w ? v : w;
    }
    return v;
}

// Inclusive prefix sum across the wave.
template <class T>
__device__ __forceinline__ T wave_incl_scan(T v) {
    const int l = __lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        T w = __shfl_up(v, o, 64);
        if (l >= o) v += w;
    }
    return v;
}

__device__ __forceinline__ uint32_t hash32(uint32_t x) { return x * 2654435761u; }

// Block-wide append of the values whose pred is set to out[], one atomic
// per call on *counter (instead of one per wave: appends from every wave
// of a large grid to one counter serialise in L2).  Every thread of the
// block must call it; lds needs NT / 64 + 1 ints.
template <int NT>
__device__ __forceinline__ void block_append(bool pred, int32_t v, int32_t *__restrict__ out,
                                             int32_t *__restrict__ counter, int32_t *lds) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t bal = __ballot(pred);
    if (lane == 0) lds[w] = __popcll(bal);
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t tot = 0;
#pragma unroll
        for (int x = 0; x < NT / 64; ++x) {
            const int32_t c = lds[x];
            lds[x] = tot;
            tot += c;
        }
        lds[NT / 64] = tot ? atomicAdd(counter, tot) : 0;
    }
    __syncthreads();
    if (pred) out[lds[NT / 64] + lds[w] + __popcll(bal & ((1ull << lane) - 1))] = v;
    __syncthreads();
}

// The same for IT candidates per thread (one atomic per NT * IT of them);
// a thread's kept values land next to each other.
template <int NT, int IT>
__device__ __forceinline__ void block_append_n(const bool (&pred)[IT], const int32_t (&v)[IT],
                                               int32_t *__restrict__ out, int32_t *__restrict__ counter,
                                               int32_t *lds) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int32_t c = 0;
#pragma unroll
    for (int i = 0; i < IT; ++i) c += pred[i];
    const int32_t incl = wave_incl_scan(c);
    if (lane == 63) lds[w] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t tot = 0;
#pragma unroll
        for (int x = 0; x < NT / 64; ++x) {
            const int32_t t = lds[x];
            lds[x] = tot;
            tot += t;
        }
        lds[NT / 64] = tot ? atomicAdd(counter, tot) : 0;
    }
    __syncthreads();
    int32_t o = lds[NT / 64] + lds[w] + incl - c;
#pragma unroll
    for (int i = 0; i < IT; ++i)
        if (pred[i]) out[o++] = v[i];
    __syncthreads();
}

}  // namespace dev
}  // namespace gk
