// k-truss and full truss decomposition on gfx950.
//
// GraphKokkos's k_truss (src/main.cpp:93-141) repeats
//     Tmp<C> = C*C;  keep entries with Tmp >= k-2 (values reset to 1)
// until an iteration removes nothing, recomputing every support each time
// and compacting on one host thread.  Here the graph stays in HBM and the
// supports are kept up to date instead of recomputed:
//     0. S<L> = L*L in SUPPORT mode: every triangle (j<k<i) adds 1 to its
//        three edges -- S<A> = A*A at A's pattern, on the oriented graph;
//     per round (same rounds as the reference's iterations):
//     1. frontier F = alive edges with S < k-2 (labelled k-1); if F is empty
//        the decomposition moves to k = min S + 3;
//     2. dS = the masked product of F (one row per edge, at its endpoint
//        with the longer row) with the alive adjacency, in PEEL mode: each
//        triangle destroyed this round takes one support off each of its
//        surviving edges;
//     3. F is removed; the adjacency is compacted on the device (per-row
//        keep counts, exclusive scan, order-preserving scatter) once
//        half of its entries are dead.
// Rounds repeat until no edge is left (decomposition) or none is removed
// (single k).  Truss numbers come back per entry of the symmetric pattern.
#include <cstddef>
#include <cstring>

#include <rocprim/device/device_scan.hpp>

#include "gk_device.h"
#include "spgemm.h"
#include "truss.h"

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

int exclusive_scan_i64(gk_ctx *ctx, const int64_t *in, int64_t *out, int64_t n) {
    size_t tmp = 0;
    GK_HIP(rocprim::exclusive_scan(nullptr, tmp, in, out, (int64_t)0, (size_t)n, rocprim::plus<int64_t>(),
                                   ctx->stream));
    void *t;
    GK_TRY(scratch(ctx, SC_CUB, tmp, &t));
    GK_HIP(rocprim::exclusive_scan(t, tmp, in, out, (int64_t)0, (size_t)n, rocprim::plus<int64_t>(), ctx->stream));
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
            int32_t mid = l + ((h - l) >> 1);
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
                    int32_t mid = l + ((h - l) >> 1);
                    if (col[mid] < (int32_t)i) l = mid + 1; else h = mid;
                }
                e = lrst[j] + (l - rst[j]);
            }
            out[p] = tl[e];
        }
    }
}

// out[p] = v[idx[p]]: per-entry values from per-edge values.
__global__ void k_gather_u32(int64_t n, const int32_t *__restrict__ idx, const uint32_t *__restrict__ v,
                             uint32_t *__restrict__ out) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
        out[p] = v[idx[p]];
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

// Device buffers of a truss run.  The first run on a context at a time
// draws them from the context's grow-only cache (the n-th allocation reuses
// the n-th cached buffer when it is large enough), so repeated runs on one
// graph allocate nothing; a nested run allocates and frees its own.
struct TrussBufs {
    gk_ctx *ctx;
    bool cached;
    size_t next = 0;
    std::vector<void *> own;
    explicit TrussBufs(gk_ctx *c) : ctx(c), cached(!c->truss_cache_busy) {
        if (cached) ctx->truss_cache_busy = true;
    }
    ~TrussBufs() {
        if (cached) ctx->truss_cache_busy = false;
        for (void *p : own) (void)hipFree(p);
    }
    template <class T>
    int alloc(T **p, size_t n) {
        const size_t bytes = sizeof(T) * (n ? n : 1);
        if (cached) {
            auto &cache = ctx->truss_cache;
            if (next == cache.size()) cache.emplace_back(nullptr, 0);
            auto &slot = cache[next++];
            if (slot.second < bytes) {
                if (slot.first) {
                    (void)hipStreamSynchronize(ctx->stream);
                    (void)hipFree(slot.first);
                    slot = {nullptr, 0};
                }
                if (hipMalloc(&slot.first, bytes) != hipSuccess) {
                    slot = {nullptr, 0};
                    set_error("hipMalloc(%zu) in truss run failed", bytes);
                    return GK_ENOMEM;
                }
                slot.second = bytes;
            }
            *p = (T *)slot.first;
            return GK_OK;
        }
        if (hipMalloc((void **)p, bytes) != hipSuccess) {
            set_error("hipMalloc(%zu) in truss run failed", bytes);
            return GK_ENOMEM;
        }
        own.push_back(*p);
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

// ---- incremental peeling ------------------------------------------------
// Edge ids are positions in L0 (the strict lower part of the input).  The
// alive adjacency Aal is the symmetric pattern with the edge id of every
// entry; it is compacted when enough of its entries are dead.

__global__ void k_edge_ends(int64_t m, const int32_t *__restrict__ lrst, const int32_t *__restrict__ lcol,
                            int32_t *__restrict__ eu, int32_t *__restrict__ ev) {
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < m; i += waves)
        for (int32_t p = lrst[i] + lane; p < lrst[i + 1]; p += 64) {
            eu[p] = (int32_t)i;
            ev[p] = lcol[p];
        }
}

// Edge id of every entry of the symmetric pattern (also to eid2 if given:
// a second copy by the same stores instead of a copy kernel afterwards).
__global__ void k_sym_eid(int64_t m, const int32_t *__restrict__ rst, const int32_t *__restrict__ col,
                          const int32_t *__restrict__ lrst, int32_t *__restrict__ eid,
                          int32_t *__restrict__ eid2 = nullptr) {
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < m; i += waves) {
        for (int32_t p = rst[i] + lane; p < rst[i + 1]; p += 64) {
            int32_t j = col[p], e;
            if (j < i) {
                e = lrst[i] + (p - rst[i]);
            } else {
                int32_t l = rst[j], h = rst[j + 1];
                while (l < h) {
                    int32_t mid = l + ((h - l) >> 1);
                    if (col[mid] < (int32_t)i) l = mid + 1; else h = mid;
                }
                e = lrst[j] + (l - rst[j]);
            }
            eid[p] = e;
            if (eid2) eid2[p] = e;
        }
    }
}

__global__ void k_iota(int64_t n, int32_t *__restrict__ a) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        a[i] = (int32_t)i;
}

// Frontier at a new threshold: alive edges with support below `thresh`,
// appended to F (their states and labels are set by the round prologue),
// and the minimum support over the alive edges (as its complement in *mn).
// With `keep`, the alive edges at or above the threshold whose support is
// below keep_hi are appended there too (the near list of the next levels).
constexpr int kScanIt = 4;   // entries per thread per pass of the list scans

// Edge states, 2 bits per edge (16 per word): 0 alive, 1 frontier (removed
// this round), 2 removed earlier.  At 2 bits the states of R-MAT scale 20's
// 15.7M edges take 3.9 MB, about an XCD's L2, where a byte each (16 MB)
// missed it on almost every random read of the peeling stream.
__device__ __forceinline__ uint32_t st_get(const uint32_t *__restrict__ st, int32_t e) {
    return (st[e >> 4] >> ((e & 15) * 2)) & 3u;
}
// alive -> frontier (00 -> 01) and frontier -> removed (01 -> 10); other
// edges of the same word may change at the same time, hence the atomics
__device__ __forceinline__ void st_to_frontier(uint32_t *st, int32_t e) {
    atomicOr(&st[e >> 4], 1u << ((e & 15) * 2));
}
__device__ __forceinline__ void st_to_removed(uint32_t *st, int32_t e) {
    atomicXor(&st[e >> 4], 3u << ((e & 15) * 2));
}

__global__ __launch_bounds__(256) void k_frontier(int64_t na, const int32_t *__restrict__ alist,
                                                  const uint32_t *__restrict__ st, const int32_t *__restrict__ sup,
                                                  int32_t thresh, int32_t *__restrict__ F, int32_t *__restrict__ nF,
                                                  uint32_t *__restrict__ mn, int32_t *__restrict__ keep,
                                                  int32_t *__restrict__ nkeep, int32_t keep_hi) {
    __shared__ int32_t lds[256 / 64 + 1];
    __shared__ uint32_t wlo[256 / 64];
    constexpr int CH = 256 * kScanIt;
    const int64_t stride = (int64_t)gridDim.x * CH;
    uint32_t lo = 0xffffffffu;
    // block-uniform trip count: every thread takes part in the appends
    for (int64_t base = (int64_t)blockIdx.x * CH; base < na; base += stride) {
        bool take[kScanIt], kp[kScanIt];
        int32_t e[kScanIt];
        int any_take = 0, any_keep = 0;
#pragma unroll
        for (int i = 0; i < kScanIt; ++i) {
            const int64_t x = base + i * 256 + threadIdx.x;
            take[i] = kp[i] = false;
            e[i] = x < na ? alist[x] : 0;
        }
#pragma unroll
        for (int i = 0; i < kScanIt; ++i) {
            const int64_t x = base + i * 256 + threadIdx.x;
            if (x < na && st_get(st, e[i]) == 0) {
                const int32_t v = sup[e[i]];
                lo = min(lo, (uint32_t)v);
                take[i] = v < thresh;
                kp[i] = keep && !take[i] && v < keep_hi;
                any_take |= take[i];
                any_keep |= kp[i];
            }
        }
        if (__syncthreads_or(any_take)) dev::block_append_n<256, kScanIt>(take, e, F, nF, lds);
        if (keep && __syncthreads_or(any_keep)) dev::block_append_n<256, kScanIt>(kp, e, keep, nkeep, lds);
    }
    lo = dev::wave_max(~lo);
    if ((threadIdx.x & 63) == 0) wlo[threadIdx.x >> 6] = lo;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t b = 0;
        for (int w = 0; w < 256 / 64; ++w) b = max(b, wlo[w]);
        if (b) atomicMax(mn, b);
    }
}

// ---- static adjacency hash for the peeling rounds ------------------------
// Vertex v's adjacency gets an open-addressing table of cap(v) = the
// smallest power of two >= max(8, 4 deg(v)) slots, (neighbour, edge id) per
// slot, built once per decomposition, and in front of it a bit filter of
// 4 cap(v) bits (a neighbour w sets bit hash(w) >> (32 - lg(4 cap))): most
// probes of the peeling stream look for a w that is not a neighbour, and a
// clear bit answers them from a 4-byte word in an array a sixteenth of the
// tables' size (90 MB against 1.4 GB at R-MAT scale 20) instead of a table
// slot.  A load factor of at most a quarter keeps the linear probe chains
// short -- a wave waits for its longest one -- and the filter at 16 bits or
// more per neighbour (R-MAT s20 115.5 -> 113.8 ms, Chung-Lu 4M 283.0 ->
// 276.6 ms against at most a half; at most an eighth: 118.5 / 279.6 ms).
// The adjacency only loses entries while peeling, so the tables stay valid:
// a removed edge is still found and then rejected by its state.

// Largest table: 2^kTableLgMax slots (vertices of degree up to 2^27); the
// peeling stream packs lg(cap) and two filter-hash bits beside a 32-bit slot
// index, which needs lg(cap) + 2 < 32.
constexpr int kTableLgMax = 29;

__global__ void k_hash_caps(int64_t m, const int32_t *__restrict__ rst, int64_t *__restrict__ cap,
                            int32_t *__restrict__ too_big) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t d = rst[i + 1] - rst[i];
        int64_t c = 0;
        if (d > 0) {
            c = 8;  // so that every table offset is a multiple of 8 (filter words at offset / 8)
            while (c < 4 * d) c <<= 1;
            if (c > (1ll << kTableLgMax)) *too_big = 1;
        }
        cap[i] = c;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) cap[m] = 0;
}

// Bit of neighbour key hash hh in the filter of a table of 2^lg slots.
__device__ __forceinline__ uint32_t filter_bit(uint32_t hh, int lg) {
    return lg + 2 >= 32 ? hh : hh >> (30 - lg);
}

// The tables filled edge by edge: edge e = (eu[e], ev[e]) goes into both
// ends' tables.  Per vertex (one wave a row) a hub's row of 10^5 entries is
// one wave's serial chain of claims while the rest of the GPU waits; per
// edge the claims of a hub's table come from every wave (Chung-Lu 4M:
// 25.4 ms per vertex).  Where a key lands depends on the order of the claims,
// what a probe finds does not.
__device__ __forceinline__ void table_put(const int64_t *__restrict__ hoff, int2 *__restrict__ tab,
                                          uint32_t *__restrict__ flt, int32_t i, int32_t w, int32_t e) {
    const int64_t h0 = hoff[i];
    const uint32_t c = (uint32_t)(hoff[i + 1] - h0);
    const int lg = 31 - __clz((int)c);
    const uint32_t hh = dev::hash32((uint32_t)w);
    uint32_t sl = hh >> (32 - lg);
    while (atomicCAS(&tab[h0 + sl].x, -1, w) != -1) sl = (sl + 1) & (c - 1);
    tab[h0 + sl].y = e;
    const uint32_t fb = filter_bit(hh, lg);
    atomicOr(&flt[(h0 >> 3) + (fb >> 5)], 1u << (fb & 31));
}

__global__ void k_hash_fill_edges(int64_t nL, const int32_t *__restrict__ eu, const int32_t *__restrict__ ev,
                                  const int64_t *__restrict__ hoff, int2 *__restrict__ tab,
                                  uint32_t *__restrict__ flt) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nL; e += (int64_t)gridDim.x * blockDim.x) {
        const int32_t u = eu[e], v = ev[e];
        table_put(hoff, tab, flt, u, v, (int32_t)e);
        table_put(hoff, tab, flt, v, u, (int32_t)e);
    }
}

// Edge id of every entry of the symmetric pattern A, flat over L's edges:
// edge e = (u, v), v < u, is entry rst[u] + e - lrst[u] of A (row u's lower
// entries are L's row u in order) and, in row v, the entry whose column is u
// (a bisection in row v).  One wave per row of A instead made a hub's row of
// 10^5 entries one wave's chain of bisections while the GPU idled
// (Chung-Lu 4M: 30.2 ms); sorting L's edges by column instead (the upper
// parts in order, no search) stalled its look-back passes beside the support
// product's kernels (R-MAT s20: first round 5 ms later).  eid2, if given,
// gets the same ids.
__global__ void k_sym_eid_edges(int64_t nL, const int32_t *__restrict__ rst, const int32_t *__restrict__ col,
                                const int32_t *__restrict__ lrst, const int32_t *__restrict__ eu,
                                const int32_t *__restrict__ ev, int32_t *__restrict__ eid,
                                int32_t *__restrict__ eid2) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nL; e += (int64_t)gridDim.x * blockDim.x) {
        const int32_t u = eu[e], v = ev[e];
        const int64_t pl = (int64_t)rst[u] + (e - lrst[u]);
        int32_t l = rst[v], h = rst[v + 1];
        while (l < h) {  // first entry of row v with column >= u
            const int32_t mid = l + ((h - l) >> 1);
            if (col[mid] < u) l = mid + 1; else h = mid;
        }
        eid[pl] = (int32_t)e;
        eid[l] = (int32_t)e;
        if (eid2) {
            eid2[pl] = (int32_t)e;
            eid2[l] = (int32_t)e;
        }
    }
}

__device__ __forceinline__ int32_t adj_probe(const int64_t *__restrict__ hoff, const int2 *__restrict__ tab,
                                             int32_t y, int32_t w) {
    const int64_t h0 = hoff[y];
    const uint32_t c = (uint32_t)(hoff[y + 1] - h0);
    const int lg = 31 - __clz((int)c);
    uint32_t s = dev::hash32((uint32_t)w) >> (32 - lg);
    for (;;) {
        const int2 kv = tab[h0 + s];
        if (kv.x == w) return kv.y;
        if (kv.x < 0) return -1;
        s = (s + 1) & (c - 1);
    }
}

// Round prologue (prep_edges).  The previous round's frontier Fp becomes removed
// (state 2); the new frontier F becomes state 1 with truss label `label`.
// Per frontier edge f = (y, x): y = endpoint with the longer current row
// (probed through its table), x = the other (its alive row is streamed;
// frs[f] = where that row starts); len[f] = current row length of x, or 0
// when f's support is 0: the
// support of a frontier edge counts its triangles that are still intact
// (every decrement of earlier rounds has landed), so such an edge closes
// no triangle this round and its row need not be streamed.
template <int P>
__device__ __forceinline__ void prep_edges(const int32_t (&f)[P], int32_t nF, const int32_t *__restrict__ F,
                                           uint32_t *__restrict__ st, uint32_t *__restrict__ truss, uint32_t label,
                                           const int32_t *__restrict__ eu, const int32_t *__restrict__ ev,
                                           const int32_t *__restrict__ sup, const int32_t *__restrict__ rst,
                                           int32_t *__restrict__ fy, int32_t *__restrict__ frs, int32_t (&len)[P]) {
    // P edges at a time, each level of loads issued for all of them before
    // any is used: edge ids -> ends and supports -> row bounds
    int32_t e[P], u[P], v[P], s[P], ru[P], ru1[P], rv[P], rv1[P];
#pragma unroll
    for (int j = 0; j < P; ++j) e[j] = f[j] < nF ? F[f[j]] : -1;
#pragma unroll
    for (int j = 0; j < P; ++j) {
        u[j] = v[j] = s[j] = 0;
        if (e[j] >= 0) {
            u[j] = eu[e[j]];
            v[j] = ev[e[j]];
            s[j] = sup[e[j]];
        }
    }
#pragma unroll
    for (int j = 0; j < P; ++j) {
        ru[j] = ru1[j] = rv[j] = rv1[j] = 0;
        if (e[j] >= 0) {
            ru[j] = rst[u[j]];
            ru1[j] = rst[u[j] + 1];
            rv[j] = rst[v[j]];
            rv1[j] = rst[v[j] + 1];
        }
    }
#pragma unroll
    for (int j = 0; j < P; ++j) {
        len[j] = 0;
        if (e[j] < 0) continue;
        st_to_frontier(st, e[j]);
        truss[e[j]] = label;
        const int32_t du = ru1[j] - ru[j], dv = rv1[j] - rv[j];
        const bool uy = du > dv || (du == dv && u[j] > v[j]);
        fy[f[j]] = uy ? u[j] : v[j];
        frs[f[j]] = uy ? rv[j] : ru[j];
        len[j] = s[j] > 0 ? (uy ? dv : du) : 0;
    }
}

// The previous frontier's edges become removed (state 2); P at a time.
template <int P>
__device__ __forceinline__ void mark_removed(int32_t nP, const int32_t *__restrict__ Fp, uint32_t *__restrict__ st,
                                             int32_t first, int32_t stride) {
    for (int32_t f0 = first; f0 < nP; f0 += P * stride) {
        int32_t e[P];
#pragma unroll
        for (int j = 0; j < P; ++j) e[j] = f0 + j * stride < nP ? Fp[f0 + j * stride] : -1;
#pragma unroll
        for (int j = 0; j < P; ++j)
            if (e[j] >= 0) st_to_removed(st, e[j]);
    }
}

// The streamed items of a round are cut into chunks of kPeelChunk (one
// wave's share per step); cstart[c] is the frontier edge owning item
// c * kPeelChunk, written by the owner of each chunk start (f with items
// [pre_f, pre_f + len) and len > 0), so the stream kernel needs no search
// over the whole frontier.  Chunks at or beyond ccap are not recorded (a
// round that long searches instead).
__device__ __forceinline__ void mark_chunks(int32_t f, int32_t pre_f, int32_t len, int32_t chunk,
                                            int32_t *__restrict__ cstart, int32_t ccap) {
    if (len <= 0) return;
    for (int32_t c = (pre_f + chunk - 1) / chunk; c < ccap && c * chunk < pre_f + len; ++c) cstart[c] = f;
}


// streamed entries per lane in flight in the peeling kernels
constexpr int kPeelU = 4;
// XCDs of the MI355X; a grid's workgroups are dispatched to them round-robin
constexpr int kXcds = 8;
// A round of at most kSpreadChunks stream chunks puts its chunks on different
// workgroups (so different CUs) in runs of kSpreadRun per XCD, instead of
// four consecutive chunks on the four waves of one workgroup: such a round
// is one step of one dependent-load chain per wave, and four waves of one CU
// queue their random loads behind each other (R-MAT s20 116.5 -> 115.3 ms,
// grid 2048^2 87.1 -> 81.1, Chung-Lu 4M 287.3 -> 283.9).  Longer rounds keep
// the four consecutive chunks together: they share owner records and table
// lines through the CU's L1 (spreading those cost R-MAT 7 ms).
constexpr int64_t kSpreadChunks = 256;
constexpr int64_t kSpreadRun = 8;
// a stream chunk whose items belong to at most kOwnerWin frontier edges
// holds their records in the wave's lanes (one each) and finds each item's
// owner there; longer runs of short rows search the prefix array instead
// (64 records instead of 17: grid 2048^2 78.3 -> 74.3 ms, R-MAT s20 111.9 -> 110.9)
constexpr int kOwnerWin = 63;

// Support update of one round, as the masked product dS<A> = F * A with the
// adjacency tables as the mask accumulator: the entries w of every frontier
// edge's streamed row are flattened over all lanes (pre = exclusive scan of
// len); w closes the triangle (y, x, w), destroyed this round; its owner --
// the frontier edge with the smallest id among its frontier edges -- takes
// one support off each of its edges that stays alive.  Edge states: 0 alive,
// 1 frontier, 2 removed earlier.
// Control block of device-driven rounds: the kernels of a batch of rounds
// read the sizes here instead of from the host.  F lives in buf[cur], the
// previous frontier and this round's crossings in buf[cur ^ 1].  Two blocks
// alternate (k_round_prep); seq numbers their writes.
struct RoundCtl {
    int32_t nP, nF, nFn, cur;
    int32_t seq;      // writes so far (the newer of the two blocks is the state)
    int32_t rounds;   // rounds run by the batch (level changes not counted)
    int32_t removed;  // frontier edges of those rounds
    int32_t adv;      // a streamed round waits for its advance (done by the next prep)
    int32_t k;        // the level: the frontier holds the edges below k - 2
    int32_t total;    // items the round streams (pre[nF])
    int32_t trans;    // this slot moves to the next level instead of streaming
    uint32_t tep;     // ... with the minimum support tagged by this prep launch
    int32_t lab;      // the last truss label given to a frontier (0: none)
};

// Moves the control block past a streamed round (or a level change): its
// frontier becomes the previous one (to be marked removed), its crossings
// (or the next level's first frontier) the next frontier.
__host__ __device__ __forceinline__ void advance_round(RoundCtl &c) {
    c.rounds += c.trans ? 0 : 1;
    c.removed += c.nF;
    c.nP = c.nF;
    c.nF = c.nFn;
    c.cur ^= 1;
    c.adv = 0;
    c.trans = 0;
}

// Control block of a decomposition in one allocation, so the host reads
// it back with one copy: the round control and the counters of the list
// scans (frontier size, complemented minimum support, near-list length,
// alive-list compaction).
struct PeelCtl {
    RoundCtl rc[2];
    unsigned long long lvmin;  // level change: prep epoch << 32 | ~(minimum support)
    int32_t knext;             // level change: the next level
    int32_t stop;  // the batch has stopped: its remaining kernels return
    int32_t err;   // a prep block's look-back wait gave up (no result)
    int32_t n;
    uint32_t mn;
    int32_t nl;
    int32_t ac;
};

// Near list: the alive edges whose support is below `hi` (plus removed
// ones, skipped when it is scanned).  The peeling kernels append an edge
// whose support drops from hi to hi - 1; hi = INT32_MAX disables it.
struct NearList {
    int32_t *list;
    int32_t *n;
    int32_t hi;
};

// A workgroup's appends to one list, staged in LDS: waves reserve room with
// an LDS atomic, and the workgroup copies the staged entries out at its
// end with one global atomic per 256 entries, instead of one global atomic
// per wave and step on the list's counter, which every wave of the grid
// shares (same-address atomics serialise in L2).  A wave whose entries do
// not fit writes them out itself and fills its part of the stage with -1.
constexpr int kStage = 1024;

struct Stage {
    int32_t n;
    int32_t v[kStage];
};

template <int N>
__device__ __forceinline__ void stage_append(const int32_t (&v)[N], Stage &S, int32_t *__restrict__ out,
                                             int32_t *__restrict__ counter) {
    int32_t c = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) c += v[i] >= 0;
    if (!__ballot(c > 0)) return;
    const int32_t incl = dev::wave_incl_scan(c);
    const int lane = threadIdx.x & 63;
    int32_t base = 0;
    if (lane == 63) base = atomicAdd(&S.n, incl);
    base = __shfl(base, 63, 64);
    const int32_t tot = __shfl(incl, 63, 64);
    int32_t o = base + incl - c;
    if (base + tot <= kStage) {
#pragma unroll
        for (int i = 0; i < N; ++i)
            if (v[i] >= 0) S.v[o++] = v[i];
        return;
    }
    for (int32_t x = base + lane; x < min(base + tot, kStage); x += 64) S.v[x] = -1;
    int32_t g = 0;
    if (lane == 63) g = atomicAdd(counter, tot);
    o = __shfl(g, 63, 64) + incl - c;
#pragma unroll
    for (int i = 0; i < N; ++i)
        if (v[i] >= 0) out[o++] = v[i];
}

// Copies a stage's entries (those >= 0) to out; every thread of the block
// calls it once, after all of the block's stage_append calls.
template <int NT>
__device__ __forceinline__ void stage_flush(Stage &S, int32_t *__restrict__ out, int32_t *__restrict__ counter,
                                            int32_t *lds) {
    __syncthreads();
    const int32_t n = min(S.n, kStage);
    for (int32_t b0 = 0; b0 < n; b0 += NT) {
        const int32_t x = b0 + (int32_t)threadIdx.x;
        const int32_t e = x < n ? S.v[x] : -1;
        const bool pred[1] = {e >= 0};
        const int32_t val[1] = {e};
        dev::block_append_n<NT, 1>(pred, val, out, counter, lds);
    }
}

// Both stages of a workgroup copied out together when each holds at most NT
// entries (the usual case): one pass, the two counters' atomics issued by
// different waves at once instead of one flush after the other.
template <int NT>
__device__ __forceinline__ void stage_flush2(Stage &A, int32_t *__restrict__ outA, int32_t *__restrict__ cntA,
                                             Stage &B, int32_t *__restrict__ outB, int32_t *__restrict__ cntB,
                                             int32_t *lds) {
    constexpr int NW = NT / 64;
    static_assert(NW >= 2, "two waves issue the two atomics");
    __syncthreads();
    const int32_t nA = min(A.n, kStage), nB = min(B.n, kStage);
    if (nA > NT || nB > NT) {
        stage_flush<NT>(A, outA, cntA, lds);
        stage_flush<NT>(B, outB, cntB, lds);
        return;
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int32_t t = (int32_t)threadIdx.x;
    const int32_t eA = t < nA ? A.v[t] : -1, eB = t < nB ? B.v[t] : -1;
    const uint64_t bA = __ballot(eA >= 0), bB = __ballot(eB >= 0);
    if (lane == 0) {
        lds[w] = __popcll(bA);
        lds[NW + 1 + w] = __popcll(bB);
    }
    __syncthreads();
    if (lane == 0 && w < 2) {  // wave 0: A's counter, wave 1: B's
        int32_t *l = lds + (w ? NW + 1 : 0);
        int32_t tot = 0;
#pragma unroll
        for (int x = 0; x < NW; ++x) {
            const int32_t c = l[x];
            l[x] = tot;
            tot += c;
        }
        l[NW] = tot ? atomicAdd(w ? cntB : cntA, tot) : 0;
    }
    __syncthreads();
    if (eA >= 0) outA[lds[NW] + lds[w] + __popcll(bA & ((1ull << lane) - 1ull))] = eA;
    if (eB >= 0) outB[lds[2 * NW + 1] + lds[NW + 1 + w] + __popcll(bB & ((1ull << lane) - 1ull))] = eB;
}

// Appends the non-negative entries of every lane's v to out, one atomic on
// *counter per wave (none when the wave has nothing).
template <int N>
__device__ __forceinline__ void wave_append(const int32_t (&v)[N], int32_t *__restrict__ out,
                                            int32_t *__restrict__ counter) {
    int32_t c = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) c += v[i] >= 0;
    if (!__ballot(c > 0)) return;
    const int32_t incl = dev::wave_incl_scan(c);
    int32_t base = 0;
    if ((threadIdx.x & 63) == 63) base = atomicAdd(counter, incl);
    base = __shfl(base, 63, 64);
    int32_t o = base + incl - c;
#pragma unroll
    for (int i = 0; i < N; ++i)
        if (v[i] >= 0) out[o++] = v[i];
}

template <int U>
__device__ __forceinline__ void k_peel_body(int32_t nF, int32_t total, const int32_t *__restrict__ pre,
                                            const int32_t *__restrict__ cstart, int32_t ccap,
                                            const int32_t *__restrict__ F,
                                            const int32_t *__restrict__ fy, const int32_t *__restrict__ frs,
                                            const int32_t *__restrict__ col, const int32_t *__restrict__ eid,
                                            const int64_t *__restrict__ hoff, const int2 *__restrict__ tab,
                                            const uint32_t *__restrict__ flt,
                                            const uint32_t *__restrict__ st, int32_t *__restrict__ sup,
                                            int32_t thresh, int32_t *__restrict__ Fn, int32_t *__restrict__ nFn,
                                            const NearList nl, unsigned long long *__restrict__ ctr) {
    constexpr int CH = 64 * U;
    // the block's staged crossings (next frontier) and near-list entries,
    // and its count of destroyed triangles
    __shared__ Stage xs, ns;
    __shared__ unsigned long long hits;
    __shared__ int32_t lds[2 * (256 / 64 + 1)];
    if (threadIdx.x == 0) {
        xs.n = ns.n = 0;
        hits = 0;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
    const int64_t nbx = gridDim.x / kXcds;  // workgroups per XCD
    int64_t wave = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if ((total + CH - 1) / CH <= kSpreadChunks && gridDim.x % kXcds == 0 &&
        (nbx * (blockDim.x >> 6)) % kSpreadRun == 0) {
        // j-th wave of XCD x (its workgroups' first waves first) -> chunk
        // (j / R) * kXcds * R + x * R + j % R: a bijection onto [0, nwaves)
        const int64_t x = blockIdx.x % kXcds, j = (int64_t)(threadIdx.x >> 6) * nbx + blockIdx.x / kXcds;
        wave = (j / kSpreadRun) * (kXcds * kSpreadRun) + x * kSpreadRun + j % kSpreadRun;
    }
    uint32_t cnt = 0;
    // Each step resolves U items per lane in stages, every stage's loads
    // issued for all U items before any is used: owner edge -> row entry
    // (column w, edge id b) and y's table bounds -> state of b and the first
    // table slot -> (rare) further probes -> state of a -> support updates.
    for (int64_t g0 = wave * CH; g0 < total; g0 += nwaves * CH) {
        const int64_t c = g0 / CH;
        int32_t f_lo, f_hi;
        if (c + 1 < ccap) {
            f_lo = cstart[c];
            f_hi = g0 + CH < total ? cstart[c + 1] : nF - 1;
        } else {  // beyond the recorded chunks: search the frontier
            const int32_t glast = (int32_t)min<int64_t>(g0 + CH, total) - 1;
            int32_t lo = 0, hi = nF - 1;
            while (lo < hi) {  // largest f with pre[f] <= g0
                const int32_t mid = lo + ((hi - lo + 1) >> 1);
                if (pre[mid] <= g0) lo = mid; else hi = mid - 1;
            }
            f_lo = lo;
            hi = nF - 1;
            while (lo < hi) {  // largest f with pre[f] <= glast
                const int32_t mid = lo + ((hi - lo + 1) >> 1);
                if (pre[mid] <= glast) lo = mid; else hi = mid - 1;
            }
            f_hi = lo;
        }
        int32_t t[U], fe[U], yy[U];
        if (f_hi - f_lo < kOwnerWin) {
            // The chunk's owners are f_lo..f_hi: lane k reads owner f_lo + k
            // whole (prefix value, edge, probed end, row start) in one
            // stage; each item's owner (the last f with pre[f] <= g) is
            // found by binary search over the lanes' prefix values, and its
            // record comes over by lane shuffle.
            static_assert(kOwnerWin < 64, "the window's search takes six steps over a wave's lanes");
            const int32_t nw = f_hi - f_lo;  // window lanes 0..nw
            const int32_t lf = f_lo + min(lane, nw);
            int32_t op = 0, oF = 0, oy = 0, orw = 0;
            if (lane <= kOwnerWin) {
                op = pre[lf];
                oF = F[lf];
                oy = fy[lf];
                orw = frs[lf];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int32_t g = (int32_t)g0 + 64 * u + lane;
                int32_t k = 0;
#pragma unroll
                for (int step = 32; step > 0; step >>= 1) {
                    const int32_t q = min(k + step, nw);
                    k = __shfl(op, q, 64) <= g ? q : k;
                }
                const int32_t pk = __shfl(op, k, 64);
                const int32_t e = __shfl(oF, k, 64), y = __shfl(oy, k, 64), r = __shfl(orw, k, 64);
                t[u] = -1;
                fe[u] = yy[u] = 0;
                if (g < total) {
                    fe[u] = e;
                    yy[u] = y;
                    t[u] = r + (g - pk);
                }
            }
        } else {
            int32_t l = f_lo;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int32_t g = (int32_t)g0 + 64 * u + lane;
                t[u] = -1;
                fe[u] = yy[u] = 0;
                if (g < total) {
                    int32_t h = f_hi;  // largest f in [l, f_hi] with pre[f] <= g (items grow with u)
                    while (l < h) {
                        const int32_t mid = l + ((h - l + 1) >> 1);
                        if (pre[mid] <= g) l = mid; else h = mid - 1;
                    }
                    fe[u] = F[l];
                    yy[u] = fy[l];
                    t[u] = frs[l] + (g - pre[l]);
                }
            }
        }
        // Per item two registers locate w in y's table and filter: hs, the
        // table index of w's first slot (y's offset h0 + slot), and lgr =
        // lg(cap) | r << 8.  Offsets are multiples of 8 and fit 32 bits, and
        // lg <= 29 (both checked by the host), so w's filter bit, slot * 4 + r
        // (r = the two hash bits below the slot's), sits in word hs / 8 at
        // bit (hs % 8) * 4 + r.
        int32_t w[U], b[U];
        uint32_t hs[U], lgr[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            w[u] = b[u] = 0;
            hs[u] = lgr[u] = 0;
            if (t[u] >= 0) {
                w[u] = col[t[u]];
                b[u] = eid[t[u]];
                const int64_t o = hoff[yy[u]];
                hs[u] = (uint32_t)o;
                lgr[u] = (uint32_t)(hoff[yy[u] + 1] - o);  // the capacity, until the next stage
            }
        }
        uint8_t s2[U];
        uint32_t fw[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            s2[u] = 2;
            fw[u] = 0;
            if (t[u] >= 0) {
                s2[u] = (uint8_t)st_get(st, b[u]);
                const int lg = 31 - __clz((int)lgr[u]);
                const uint32_t hh = dev::hash32((uint32_t)w[u]);
                hs[u] += hh >> (32 - lg);
                lgr[u] = (uint32_t)lg | (((hh >> (30 - lg)) & 3u) << 8);
                fw[u] = flt[hs[u] >> 3];
            }
        }
        // y's table is read only where the filter bit of w is set (w may be
        // a neighbour of y) and the streamed edge can close a triangle
        bool look[U];
        int2 kv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            look[u] = t[u] >= 0 && s2[u] != 2 && !(s2[u] == 1 && b[u] < fe[u]) &&
                      ((fw[u] >> (((hs[u] & 7u) << 2) | (lgr[u] >> 8))) & 1u);
            kv[u] = look[u] ? tab[hs[u]] : make_int2(-1, -1);
        }
        int32_t a[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            a[u] = -1;
            if (!look[u]) continue;
            if (kv[u].x == w[u]) {
                a[u] = kv[u].y;
                continue;
            }
            // further slots (rare): y's table starts at h0 = hs - slot
            const uint32_t lg = lgr[u] & 0xffu, cap = 1u << lg;
            uint32_t slot = dev::hash32((uint32_t)w[u]) >> (32 - lg);
            const uint32_t h0 = hs[u] - slot;
            while (kv[u].x >= 0) {  // y's table: (neighbour, edge id), -1 marks an empty slot
                slot = (slot + 1) & (cap - 1);
                kv[u] = tab[h0 + slot];
                if (kv[u].x == w[u]) {
                    a[u] = kv[u].y;
                    break;
                }
            }
        }
        uint8_t s1[U];
#pragma unroll
        for (int u = 0; u < U; ++u) s1[u] = a[u] >= 0 ? (uint8_t)st_get(st, a[u]) : 2;
        // edges leaving the threshold / entering the near list (-1: none)
        int32_t xf[2 * U], xn[2 * U];
#pragma unroll
        for (int u = 0; u < 2 * U; ++u) xf[u] = xn[u] = -1;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (a[u] < 0 || s1[u] == 2 || (s1[u] == 1 && a[u] < fe[u])) continue;
            ++cnt;
            // an alive edge whose support falls below the threshold right
            // now (old value == thresh; supports drop by one at a time)
            // joins the next round's frontier; one falling below the near
            // list's bound joins the near list
            if (s1[u] == 0) {
                const int32_t old = atomicSub(&sup[a[u]], 1);
                if (old == thresh) xf[2 * u] = a[u];
                else if (old == nl.hi) xn[2 * u] = a[u];
            }
            if (s2[u] == 0) {
                const int32_t old = atomicSub(&sup[b[u]], 1);
                if (old == thresh) xf[2 * u + 1] = b[u];
                else if (old == nl.hi) xn[2 * u + 1] = b[u];
            }
        }
        stage_append(xf, xs, Fn, nFn);
        stage_append(xn, ns, nl.list, nl.n);
    }
    unsigned long long s = dev::wave_sum((unsigned long long)cnt);
    if (lane == 0 && s) atomicAdd(&hits, s);
    stage_flush2<256>(xs, Fn, nFn, ns, nl.list, nl.n, lds);
    if (threadIdx.x == 0) {
        if (hits) atomicAdd(&ctr[1], hits);
        if (blockIdx.x == 0) atomicAdd(&ctr[0], (unsigned long long)total);
    }
}

// streamed items per wave step of the peeling kernels
constexpr int kPeelChunk = 64 * kPeelU;

// Device-driven rounds, two kernels each: k_round_prep, then
// k_dround_stream.  Round r's prep reads the control block C[r & 1] (the
// previous round's state, its advance pending when adv is set) and writes
// the advanced state to C[(r + 1) & 1], which the round's stream reads and
// whose nFn counts the stream's crossings; so no block is written while a
// kernel of the same launch may still read it.  The prep marks the previous
// frontier removed and the new one frontier (labels, streamed rows, their
// lengths) and scans the lengths into pre in the same launch: each workgroup
// owns a contiguous piece of the frontier and finds the items before it by
// decoupled look-back over the workgroups before it (workgroups are
// dispatched in index order, and each waits only on lower ones).
//
// An empty frontier ends a level.  In a decomposition whose near list is
// valid, the slot changes level on the device: when a level ends every
// alive edge has support >= k - 2, so the next level k' = max(k + 1, s + 3)
// (s = the minimum support over the alive edges) starts from the alive
// edges below k' - 2 = s + 1, i.e. those at support s -- all of them in the
// near list, whose alive entries are exactly the edges below its bound.  The
// prep takes s over the near list, the stream collects the edges at s as
// the next frontier.  Otherwise (single k, no valid near list, or a near
// list with no alive entry) the stop word ends the batch and the host scans.
constexpr int kPrepThreads = 1024;
// pinned staging of a decomposition: the tables' size, scan counters and a
// PeelCtl copy
constexpr int kPinnedBytes = 256;
constexpr int kPrepPer = 4;        // frontier edges per thread per pass
constexpr int kPrepBlocksMax = 128;
// look-back polls before a prep workgroup gives up (about a second; a
// workgroup that gives up sets the error word, and the host fails the run)
constexpr uint32_t kLookbackSpins = 1u << 20;

// Look-back word of a prep workgroup: epoch << 33 | inclusive << 32 | items
// (its own, or with every workgroup before it when inclusive).  The epoch
// is the launch's number in the decomposition, so words of earlier rounds
// read as not ready.
__device__ __forceinline__ unsigned long long lb_word(uint32_t epoch, bool incl, uint32_t v) {
    return ((unsigned long long)epoch << 33) | ((unsigned long long)incl << 32) | v;
}

// What a batch's kernels share besides the control blocks.
struct PeelDev {
    int32_t *stop, *err;
    unsigned long long *lvmin;
    int32_t *knext;
    unsigned long long *lb;  // look-back words, one per prep workgroup
    const int32_t *nl;       // near list (entries *nl_n) for level changes
    const int32_t *nl_n;
    int32_t levels;          // the device may change level
    uint32_t fixed_label;    // single k: every removed edge gets this label (else k - 1)
};

__global__ __launch_bounds__(kPrepThreads) void k_round_prep(
    const RoundCtl *__restrict__ ci, RoundCtl *__restrict__ co, const PeelDev D, uint32_t epoch,
    int32_t *__restrict__ buf0, int32_t *__restrict__ buf1, uint32_t *__restrict__ st, uint32_t *__restrict__ truss,
    const int32_t *__restrict__ eu, const int32_t *__restrict__ ev, const int32_t *__restrict__ sup,
    const int32_t *__restrict__ rst, int32_t *__restrict__ fy, int32_t *__restrict__ fx,
    int32_t *__restrict__ len, int32_t *__restrict__ pre, int32_t *__restrict__ cstart, int32_t ccap) {
    constexpr int NW = kPrepThreads / 64;
    __shared__ int32_t ws[NW];
    __shared__ uint32_t wlo[NW];
    __shared__ int32_t excl_s;
    // the stop word and the control block are loaded together (one round
    // trip; the test reads both, and nF is never negative); a stopped
    // batch's control block is still a valid one
    const int32_t stopped = *D.stop;
    RoundCtl c = *ci;
    if (stopped | (c.nF < 0)) return;
    if (c.adv) {
        const bool lev = c.trans && c.nFn > 0;  // the slot before collected a new level's frontier
        advance_round(c);
        if (lev) c.k = *D.knext;
    }
    c.seq += 1;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int32_t nF = c.nF, nP = c.nP;
    const int32_t *F = c.cur ? buf1 : buf0;
    const int32_t *Fp = c.cur ? buf0 : buf1;
    // The previous frontier becomes removed, workgroup b taking the entries
    // [p0, p1): by all its threads where it has no frontier edges of its own,
    // else by waves 1.. while wave 0 waits in the look-back (the new and the
    // previous frontier are disjoint, and the stream reads the states only
    // after this launch), so the marking is off the launch's load chain.
    const int32_t pc = (nP + (int32_t)gridDim.x - 1) / (int32_t)gridDim.x;
    const int32_t p0 = min(nP, (int32_t)blockIdx.x * pc), p1 = min(nP, p0 + pc);
    c.nP = 0;
    if (nF == 0) {
        mark_removed<4>(p1, Fp, st, p0 + tid, kPrepThreads);
        if (D.levels) {
            // minimum support over the near list's alive entries
            uint32_t lo = 0xffffffffu;
            const int32_t nn = *D.nl_n;
            for (int32_t x = blockIdx.x * kPrepThreads + tid; x < nn; x += gridDim.x * kPrepThreads) {
                const int32_t e = D.nl[x];
                if (st_get(st, e) == 0) lo = min(lo, (uint32_t)sup[e]);
            }
            lo = dev::wave_max(~lo);
            if (lane == 0) wlo[w] = lo;
            __syncthreads();
            if (tid == 0) {
                uint32_t m = 0;
                for (int x = 0; x < NW; ++x) m = max(m, wlo[x]);
                if (m) atomicMax(D.lvmin, ((unsigned long long)epoch << 32) | m);
            }
            if (blockIdx.x == 0 && tid == 0) {
                c.trans = 1;
                c.tep = epoch;
                c.nFn = 0;
                c.adv = 1;
                c.total = 0;
                *co = c;
            }
        } else if (blockIdx.x == 0 && tid == 0) {
            *co = c;
            *D.stop = 1;
        }
        return;
    }
    const uint32_t label = D.fixed_label ? D.fixed_label : (uint32_t)(c.k - 1);
    const int32_t chunk = (nF + gridDim.x - 1) / gridDim.x;
    const int32_t f0 = min(nF, (int32_t)blockIdx.x * chunk), f1 = min(nF, f0 + chunk);
    // workgroups past the frontier's end have nothing more to do: every
    // workgroup after them is empty too, so none waits on their look-back
    // words (a frontier of fewer edges than workgroups ends the launch early)
    if (f0 >= f1) {
        mark_removed<4>(p1, Fp, st, p0 + tid, kPrepThreads);
        return;
    }
    // pass 1: marks, labels, streamed rows and their lengths; the piece's sum
    int32_t sum = 0;
    for (int32_t b = f0 + tid; b < f1; b += kPrepPer * kPrepThreads) {
        int32_t fi[kPrepPer], l[kPrepPer];
#pragma unroll
        for (int j = 0; j < kPrepPer; ++j) fi[j] = b + j * kPrepThreads < f1 ? b + j * kPrepThreads : nF;
        prep_edges<kPrepPer>(fi, nF, F, st, truss, label, eu, ev, sup, rst, fy, fx, l);
#pragma unroll
        for (int j = 0; j < kPrepPer; ++j)
            if (fi[j] < nF) {
                len[fi[j]] = l[j];
                sum += l[j];
            }
    }
    sum = dev::wave_sum(sum);
    if (lane == 0) ws[w] = sum;
    __syncthreads();
    // look-back (wave 0): items of the workgroups before this one
    if (w == 0) {
        int32_t agg = 0;
#pragma unroll
        for (int x = 0; x < NW; ++x) agg += ws[x];
        uint32_t excl = 0;
        const int32_t bi = blockIdx.x;
        if (lane == 0) __hip_atomic_store(&D.lb[bi], lb_word(epoch, bi == 0, (uint32_t)agg), __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
        int32_t look = bi - 1;  // highest workgroup not yet added
        uint32_t spins = 0;
        while (look >= 0) {
            const int32_t p = look - lane;
            // workgroups below 0 read as an inclusive zero, so a window
            // always ends at workgroup 0 (whose word is inclusive)
            const unsigned long long v = p >= 0 ? __hip_atomic_load(&D.lb[p], __ATOMIC_RELAXED,
                                                                    __HIP_MEMORY_SCOPE_AGENT)
                                                : lb_word(epoch, true, 0);
            const bool ready = (uint32_t)(v >> 33) == epoch;
            const uint64_t incl = __ballot(ready && ((v >> 32) & 1ull));
            const int first = incl ? __builtin_ctzll(incl) : 64;  // nearest inclusive word
            const uint64_t need = first >= 63 ? ~0ull : ((2ull << first) - 1ull);
            if (__ballot(!ready) & need) {
                if (++spins > kLookbackSpins) {
                    if (lane == 0) atomicExch(D.err, 1);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            excl += dev::wave_sum(lane <= first ? (uint32_t)v : 0u);
            if (first < 64) break;
            look -= 64;
        }
        if (bi > 0 && lane == 0)
            __hip_atomic_store(&D.lb[bi], lb_word(epoch, true, excl + (uint32_t)agg), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        if (lane == 0) excl_s = (int32_t)excl;
    } else {
        mark_removed<4>(p1, Fp, st, p0 + tid - 64, kPrepThreads - 64);
    }
    __syncthreads();
    // pass 2: the piece's lengths scanned from its offset into pre, and the
    // owners of the stream chunks that start in it
    int32_t carry = excl_s;
    for (int32_t c0 = f0; c0 < f1; c0 += kPrepThreads) {
        const int32_t f = c0 + tid;
        const int32_t l = f < f1 ? len[f] : 0;
        const int32_t incl = dev::wave_incl_scan(l);
        if (lane == 63) ws[w] = incl;
        __syncthreads();
        int32_t before = 0, tot = 0;
#pragma unroll
        for (int x = 0; x < NW; ++x) {
            const int32_t v = ws[x];
            before += x < w ? v : 0;
            tot += v;
        }
        if (f < f1) {
            const int32_t ex = carry + before + incl - l;
            pre[f] = ex;
            mark_chunks(f, ex, l, kPeelChunk, cstart, ccap);
        }
        carry += tot;
        __syncthreads();
    }
    // the workgroup holding the frontier's end knows the round's items and
    // writes the control block for the stream
    if (f1 == nF && f0 < f1 && tid == 0) {
        pre[nF] = carry;
        c.total = carry;
        c.nFn = 0;
        c.adv = 1;  // the stream kernel that follows runs this round
        c.lab = (int32_t)label;
        *co = c;
    }
}

// Eight waves per SIMD (62 VGPRs; the scalar registers over 100 that would
// cap it at seven are spilled to vector lanes): most rounds are chains of
// dependent loads, and the long ones are bound by the waves in flight
// (R-MAT scale 20: 123.5 -> 117.1 ms against seven waves).
template <int U>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_dround_stream(const RoundCtl *__restrict__ ctl, const PeelDev D,
                                                       int32_t *__restrict__ buf0, int32_t *__restrict__ buf1,
                                                       const int32_t *__restrict__ pre,
                                                       const int32_t *__restrict__ fy, const int32_t *__restrict__ fx,
                                                       const int32_t *__restrict__ cstart, int32_t ccap,
                                                       const int32_t *__restrict__ col,
                                                       const int32_t *__restrict__ eid,
                                                       const int64_t *__restrict__ hoff, const int2 *__restrict__ tab,
                                                       const uint32_t *__restrict__ flt,
                                                       const uint32_t *__restrict__ st, int32_t *__restrict__ sup,
                                                       int32_t *__restrict__ nFn, const NearList nl,
                                                       unsigned long long *__restrict__ ctr) {
    const int32_t stopped = *D.stop;
    const RoundCtl c = *ctl;
    if (stopped | (c.nF < 0)) return;  // both loads at once, as in k_round_prep
    int32_t *Fn = c.cur ? buf0 : buf1;
    if (c.trans) {
        // level change: the alive near-list edges at the minimum support s
        // (tagged by the prep of this slot) are the next level's frontier
        const unsigned long long wv = *D.lvmin;
        if ((uint32_t)(wv >> 32) != c.tep) {  // no alive entry: the host scans
            if (blockIdx.x == 0 && threadIdx.x == 0) *D.stop = 1;
            return;
        }
        const int32_t s = (int32_t)~(uint32_t)wv;
        if (blockIdx.x == 0 && threadIdx.x == 0) *D.knext = max(c.k + 1, s + 3);
        __shared__ int32_t lds[256 / 64 + 1];
        const int32_t nn = *D.nl_n;
        constexpr int CH = 256 * kScanIt;
        for (int32_t base = blockIdx.x * CH; base < nn; base += gridDim.x * CH) {
            bool take[kScanIt];
            int32_t e[kScanIt];
#pragma unroll
            for (int i = 0; i < kScanIt; ++i) {
                const int32_t x = base + i * 256 + (int32_t)threadIdx.x;
                e[i] = x < nn ? D.nl[x] : 0;
            }
            int any = 0;
#pragma unroll
            for (int i = 0; i < kScanIt; ++i) {
                const int32_t x = base + i * 256 + (int32_t)threadIdx.x;
                take[i] = x < nn && st_get(st, e[i]) == 0 && sup[e[i]] == s;
                any |= take[i];
            }
            if (__syncthreads_or(any)) dev::block_append_n<256, kScanIt>(take, e, Fn, nFn, lds);
        }
        return;
    }
    // the crossings go to buf[cur ^ 1], counted in ctl->nFn
    k_peel_body<U>(c.nF, c.total, pre, cstart, ccap, c.cur ? buf1 : buf0, fy, fx, col, eid, hoff, tab, flt, st,
                   sup, c.k - 2, Fn, nFn, nl, ctr);
}


__global__ void k_mark_dead(int32_t nF, const int32_t *__restrict__ F, uint32_t *__restrict__ st) {
    for (int32_t f = blockIdx.x * blockDim.x + threadIdx.x; f < nF; f += gridDim.x * blockDim.x) st_to_removed(st, F[f]);
}

// Compaction of the alive adjacency (rows stay sorted).
// Adjacency compaction: rows of up to kLongRow entries one wave each, longer
// ones (a power-law graph's hubs: 10^5 entries would be one wave's chain of
// thousands of dependent steps) one 1024-thread workgroup each.
constexpr int32_t kLongRow = 4096;
constexpr int kLongThreads = 1024;

__global__ void k_alive_count(int64_t m, const int32_t *__restrict__ rst, const int32_t *__restrict__ eid,
                              const uint32_t *__restrict__ st, int32_t *__restrict__ cnt) {
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < m; i += waves) {
        if (rst[i + 1] - rst[i] > kLongRow) continue;
        int32_t c = 0;
        for (int32_t p = rst[i] + lane; p < rst[i + 1]; p += 64) c += st_get(st, eid[p]) != 2;
        c = dev::wave_sum(c);
        if (lane == 0) cnt[i] = c;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) cnt[m] = 0;
}

__global__ void k_alive_scatter(int64_t m, const int32_t *__restrict__ rst, const int32_t *__restrict__ col,
                                const int32_t *__restrict__ eid, const uint32_t *__restrict__ st,
                                const int32_t *__restrict__ nrst, int32_t *__restrict__ ncol,
                                int32_t *__restrict__ neid) {
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < m; i += waves) {
        if (rst[i + 1] - rst[i] > kLongRow) continue;
        int32_t out = nrst[i];
        for (int32_t base = rst[i]; base < rst[i + 1]; base += 64) {
            int32_t p = base + lane;
            bool k = p < rst[i + 1] && st_get(st, eid[p]) != 2;
            uint64_t bal = __ballot(k);
            if (k) {
                int32_t o = out + __popcll(bal & ((1ull << lane) - 1));
                ncol[o] = col[p];
                neid[o] = eid[p];
            }
            out += __popcll(bal);
        }
    }
}

// The rows longer than kLongRow, one workgroup each: workgroup b reads the
// row pointers of rows b, b + grid, b + 2 grid, ... (with degree-ordered
// ids the long rows are the first ones, so they spread over the
// workgroups), lists its long rows in LDS, and counts (SCATTER false) or
// compacts (SCATTER true, tile by tile in order) them.
template <bool SCATTER>
__global__ __launch_bounds__(kLongThreads) void k_alive_long(int64_t m, const int32_t *__restrict__ rst,
                                                             const int32_t *__restrict__ col,
                                                             const int32_t *__restrict__ eid,
                                                             const uint32_t *__restrict__ st,
                                                             int32_t *__restrict__ cnt,
                                                             const int32_t *__restrict__ nrst,
                                                             int32_t *__restrict__ ncol, int32_t *__restrict__ neid) {
    constexpr int NW = kLongThreads / 64;
    __shared__ int32_t rows[kLongThreads];
    __shared__ int32_t nrows;
    __shared__ int32_t ws[NW];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    for (int64_t k0 = 0; (int64_t)blockIdx.x + k0 * gridDim.x < m; k0 += kLongThreads) {
        if (tid == 0) nrows = 0;
        __syncthreads();
        const int64_t i = (int64_t)blockIdx.x + (k0 + tid) * gridDim.x;
        if (i < m && rst[i + 1] - rst[i] > kLongRow) rows[atomicAdd(&nrows, 1)] = (int32_t)i;
        __syncthreads();
        const int32_t nr = nrows;
        for (int32_t r = 0; r < nr; ++r) {
            const int32_t row = rows[r];
            const int32_t b0 = rst[row], b1 = rst[row + 1];
            int32_t out = SCATTER ? nrst[row] : 0;
            for (int32_t base = b0; base < b1; base += kLongThreads) {
                const int32_t p = base + tid;
                const bool k = p < b1 && st_get(st, eid[p]) != 2;
                const uint64_t bal = __ballot(k);
                if (lane == 0) ws[w] = __popcll(bal);
                __syncthreads();
                int32_t before = 0, tot = 0;
#pragma unroll
                for (int x = 0; x < NW; ++x) {
                    const int32_t v = ws[x];
                    before += x < w ? v : 0;
                    tot += v;
                }
                if (SCATTER && k) {
                    const int32_t o = out + before + __popcll(bal & ((1ull << lane) - 1ull));
                    ncol[o] = col[p];
                    neid[o] = eid[p];
                }
                out += tot;
                __syncthreads();
            }
            if (!SCATTER && tid == 0) cnt[row] = out;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_alist_compact(int64_t na, const int32_t *__restrict__ alist,
                                                       const uint32_t *__restrict__ st, int32_t *__restrict__ out,
                                                       int32_t *__restrict__ nout) {
    __shared__ int32_t lds[256 / 64 + 1];
    constexpr int CH = 256 * kScanIt;
    const int64_t stride = (int64_t)gridDim.x * CH;
    for (int64_t base = (int64_t)blockIdx.x * CH; base < na; base += stride) {
        bool k[kScanIt];
        int32_t e[kScanIt];
#pragma unroll
        for (int i = 0; i < kScanIt; ++i) {
            const int64_t x = base + i * 256 + threadIdx.x;
            e[i] = x < na ? alist[x] : 0;
        }
#pragma unroll
        for (int i = 0; i < kScanIt; ++i) k[i] = base + i * 256 + threadIdx.x < na && st_get(st, e[i]) == 0;
        dev::block_append_n<256, kScanIt>(k, e, out, nout, lds);
    }
}

// Blocks for a scan of n list entries (kScanIt per thread).
int grid_scan(gk_ctx *ctx, int64_t n) {
    return (int)std::max<int64_t>(1, std::min<int64_t>((n + 256 * kScanIt - 1) / (256 * kScanIt),
                                                       (int64_t)ctx->num_cus * 8));
}

int grid_items(gk_ctx *ctx, int64_t n) {
    return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, (int64_t)ctx->num_cus * 16));
}

// Truss numbers of every L0 edge by incremental peeling.  Round:
//   F = alive edges with support < k-2 (labelled k-1);  none -> k = min + 3;
//   dS = F * A masked by A (k_dround_stream): one support off each
//        surviving edge of every triangle destroyed this round;
//   F -> removed; compact the adjacency once half of it is dead.
// Rounds run in device-driven batches (sizes in the RoundCtl blocks), each
// round a prep (marks + scan) and a stream launch; a batch stops at the end
// of a level, and the host checks in once per batch.
// fixed_k > 0 runs the single-k semantics instead (k_truss, src/main.cpp:93-141):
// the threshold stays k-2, removed edges are labelled 1, and the run stops at
// the first round that removes nothing; the returned count then includes that
// last round, as the reference's iteration count does.
int truss_peel(gk_ctx *ctx, const gk_dcsr *A, TrussBufs &B, const Lgraph &L0, uint32_t *tl, const gk_opts &o,
               int64_t *rounds_out, int32_t *kmax_out, int fixed_k = 0, int32_t **sym_eid_out = nullptr) {
    const int64_t m = A->m, nL = L0.nz;
    int32_t *sup, *eu, *ev, *alist, *alist2, *F, *Fn, *fy, *fx, *flen, *fpre, *cnt;
    int32_t *arst[2], *acol[2], *aeid[2];
    int64_t *caps, *hoff;
    int2 *tab;
    uint32_t *st;  // edge states, 2 bits each
    unsigned long long *dctr;  // [0] items streamed, [1] triangles destroyed
    PeelCtl *dpc;
    GK_TRY(B.alloc(&sup, nL));
    GK_TRY(B.alloc(&eu, nL));
    GK_TRY(B.alloc(&ev, nL));
    GK_TRY(B.alloc(&alist, nL));
    GK_TRY(B.alloc(&alist2, nL));
    GK_TRY(B.alloc(&F, nL));
    GK_TRY(B.alloc(&Fn, nL));
    GK_TRY(B.alloc(&fy, nL));
    GK_TRY(B.alloc(&fx, nL));
    GK_TRY(B.alloc(&flen, nL + 1));
    GK_TRY(B.alloc(&fpre, nL + 1));
    // owner of every chunk start, for rounds of up to about A.nz items
    // (longer rounds search the frontier instead)
    const int32_t ccap = (int32_t)std::min<int64_t>(A->nz / kPeelChunk + 2, INT32_MAX);
    int32_t *cstart;
    GK_TRY(B.alloc(&cstart, ccap));
    const int64_t st_words = (nL + 15) / 16;
    GK_TRY(B.alloc(&st, st_words));
    GK_TRY(B.alloc(&cnt, m + 1));
    GK_TRY(B.alloc(&caps, m + 1));
    GK_TRY(B.alloc(&hoff, m + 1));
    GK_TRY(B.alloc(&dctr, 2));
    GK_TRY(B.alloc(&dpc, 1));
    int32_t *dn = &dpc->n, *dnl = &dpc->nl, *dac = &dpc->ac;
    uint32_t *dmin = &dpc->mn;
    int32_t *NL, *NL2;
    GK_TRY(B.alloc(&NL, nL));
    GK_TRY(B.alloc(&NL2, nL));
    for (int b = 0; b < 2; ++b) {
        GK_TRY(B.alloc(&arst[b], m + 1));
        GK_TRY(B.alloc(&acol[b], A->nz));
        GK_TRY(B.alloc(&aeid[b], A->nz));
    }
    const int gw = waves_grid(ctx, m);
    // adjacency tables: their sizes first (the table is allocated from them)
    GK_HIP(hipMemsetAsync(&dpc->ac, 0, 4, ctx->stream));
    k_hash_caps<<<grid_items(ctx, m), 256, 0, ctx->stream>>>(m, A->rst, caps, &dpc->ac);
    GK_TRY(exclusive_scan_i64(ctx, caps, hoff, m + 1));
    int64_t *h64;
    GK_TRY(pinned(ctx, kPinnedBytes, (void **)&h64));
    GK_HIP(hipMemcpyAsync(h64, hoff + m, 8, hipMemcpyDeviceToHost, ctx->stream));
    GK_HIP(hipMemcpyAsync(h64 + 1, &dpc->ac, 4, hipMemcpyDeviceToHost, ctx->stream));
    GK_HIP(hipStreamSynchronize(ctx->stream));
    const int64_t tab_n = h64[0];
    // the stream indexes the tables with 32-bit slot numbers
    if (*(const int32_t *)(h64 + 1) || tab_n > (int64_t)UINT32_MAX) {
        set_error("truss peeling: the adjacency tables need a vertex of degree at most 2^27 and at most 2^32 "
                  "slots in all");
        return GK_EINVAL;
    }
    GK_TRY(B.alloc(&tab, tab_n));
    uint32_t *tflt;  // the tables' bit filters: tab_n / 8 words (every cap is a multiple of 8)
    GK_TRY(B.alloc(&tflt, tab_n / 8));
    if (sym_eid_out) GK_TRY(B.alloc(sym_eid_out, A->nz));
    // The set-up that does not need the supports (edge ends, the adjacency
    // with edge ids, the tables, the lists) runs on the side stream, beside
    // the support product on the others; it writes none of the product's
    // buffers.
    GK_HIP(hipEventRecord(ctx->side_fork, ctx->stream));
    GK_HIP(hipStreamWaitEvent(ctx->side, ctx->side_fork, 0));
    {
        hipStream_t s = ctx->side;
        const int ge = grid_items(ctx, nL);
        k_edge_ends<<<gw, 256, 0, s>>>(m, L0.rst, L0.col, eu, ev);
        GK_HIP(hipMemcpyAsync(arst[0], A->rst, 4 * (m + 1), hipMemcpyDeviceToDevice, s));
        GK_HIP(hipMemcpyAsync(acol[0], A->col, 4 * (size_t)A->nz, hipMemcpyDeviceToDevice, s));
        k_iota<<<ge, 256, 0, s>>>(nL, alist);
        GK_HIP(hipMemsetAsync(st, 0, 4 * (size_t)std::max<int64_t>(st_words, 1), s));
        GK_HIP(hipMemsetAsync(dctr, 0, 16, s));
        GK_HIP(hipMemsetAsync(tab, 0xff, sizeof(int2) * (size_t)std::max<int64_t>(tab_n, 1), s));
        GK_HIP(hipMemsetAsync(tflt, 0, 4 * (size_t)std::max<int64_t>(tab_n / 8, 1), s));
        k_hash_fill_edges<<<ge, 256, 0, s>>>(nL, eu, ev, hoff, tab, tflt);
        k_sym_eid_edges<<<ge, 256, 0, s>>>(nL, A->rst, A->col, L0.rst, eu, ev, aeid[0],
                                           sym_eid_out ? *sym_eid_out : nullptr);
        GK_TRY(launch_check("truss_peel setup"));
        GK_HIP(hipEventRecord(ctx->side_join, s));
    }
    // supports of the whole graph: C<L0> = L0*L0 in SUPPORT mode
    {
        gk_dcsr Ld = view(ctx, m, m, nL, L0.rst, L0.col, nullptr);
        ProductInfo pi;
        const int rc = masked_product(ctx, &Ld, &Ld, &Ld, MODE_SUPPORT, sup, nullptr, nullptr, o, &pi);
        GK_HIP(hipStreamWaitEvent(ctx->stream, ctx->side_join, 0));
        GK_TRY(rc);
        record_timing(ctx, pi, false);
    }
    int cur = 0;                 // adjacency buffer in use
    int64_t aal_nz = A->nz;      // entries in the current adjacency
    int64_t na = nL;             // entries in the alive list
    int64_t alive = nL, dead_pending = 0;
    int64_t rounds = 0;
    int k = fixed_k > 0 ? fixed_k : 3;
    int32_t kbest = nL ? 2 : 0;
    int32_t *h = (int32_t *)(h64 + 2);
    static_assert(sizeof(PeelCtl) <= kPinnedBytes - 32, "control block fits the pinned staging");
    PeelCtl *hpc = (PeelCtl *)(h64 + 4);
    // Peel-stream timing: HIP events around every kPeelTimeStride-th stream
    // launch only.  A timed event drains the queue; around every launch
    // (3,677 on R-MAT scale 20) they cost 30 ms of a 233 ms decomposition.
    // The sampled durations, scaled by launches / sampled launches, give
    // the stream kernels' total.
    constexpr int64_t kPeelTimeStride = 8;
    std::vector<hipEvent_t> &ev_pool = ctx->events;
    size_t ev_used = 0;
    int64_t stream_launches = 0;
    auto ev_pair = [&](size_t *at) -> int {
        while (ev_used + 2 > ev_pool.size()) {
            hipEvent_t e;
            GK_HIP(hipEventCreate(&e));
            ev_pool.push_back(e);
        }
        *at = ev_used;
        ev_used += 2;
        return GK_OK;
    };
    // events before and after a stream launch, on the sampled launches
    size_t ev_cur = 0;
    auto stream_begin = [&]() -> int {
        if (stream_launches++ % kPeelTimeStride) return GK_OK;
        GK_TRY(ev_pair(&ev_cur));
        GK_HIP(hipEventRecord(ev_pool[ev_cur], ctx->stream));
        return GK_OK;
    };
    auto stream_end = [&]() -> int {
        if ((stream_launches - 1) % kPeelTimeStride) return GK_OK;
        GK_HIP(hipEventRecord(ev_pool[ev_cur + 1], ctx->stream));
        return GK_OK;
    };
    const int gs = (int)std::max<int64_t>(1, (int64_t)ctx->num_cus * 8);
    int32_t *fb[2] = {F, Fn};    // frontier buffers: F in fb[fc], previous / crossings in fb[fc ^ 1]
    int fc = 0;
    int32_t nP = 0;              // previous frontier, still to be marked removed
    int32_t nF = 0;              // current frontier
    bool scan_needed = true;     // the frontier comes from a scan
    bool started = false;
    // Device-driven rounds per host check, adapted to the cascades: a batch
    // whose rounds all ran doubles the next one; a batch cut short (the
    // level ended) sizes the next one from the rounds it ran, so few
    // launches are wasted on stopped rounds.
    int batch = 8;
    auto next_batch = [](int size, int32_t ran, int cap) {
        return ran >= size ? std::min(2 * size, cap) : std::max(2, std::min<int>(ran + ran / 2 + 1, cap));
    };
    const int nprep = (int)std::min<int64_t>(kPrepBlocksMax, (int64_t)ctx->num_cus);
    unsigned long long *lb;  // look-back words of the prep workgroups
    GK_TRY(B.alloc(&lb, kPrepBlocksMax));
    GK_HIP(hipMemsetAsync(lb, 0, 8 * kPrepBlocksMax, ctx->stream));
    uint32_t epoch = 0;
    // Near list (NearList): alive edges with support below nl_hi, built by a
    // full scan of the alive list and kept complete by the peeling kernels.
    // While a level's threshold stays below nl_hi its frontier comes from
    // the near list, compacted as it is scanned.  The window nl_hi - thresh
    // adapts so the list holds about an eighth of the alive list.
    bool nl_valid = false;
    int32_t nl_hi = INT32_MAX;
    int64_t nNL = 0, nl_window = 8;
    auto near_list = [&]() { return NearList{NL, dnl, nl_valid ? nl_hi : INT32_MAX}; };
    // one scan at threshold thr: nF, the minimum support over the scanned
    // edges (h[1], complemented) and the near list's new length
    auto scan = [&](bool full, int32_t thr) -> int {
        GK_HIP(hipMemsetAsync(dn, 0, 12, ctx->stream));  // n, mn, nl
        if (full) {
            const int32_t hi = fixed_k > 0 ? thr : (int32_t)std::min<int64_t>((int64_t)thr + nl_window, INT32_MAX - 1);
            k_frontier<<<grid_scan(ctx, na), 256, 0, ctx->stream>>>(na, alist, st, sup, thr, fb[fc], dn, dmin,
                                                                     fixed_k > 0 ? nullptr : NL, dnl, hi);
            nl_hi = hi;
        } else {
            k_frontier<<<grid_scan(ctx, nNL), 256, 0, ctx->stream>>>(nNL, NL, st, sup, thr, fb[fc], dn, dmin, NL2,
                                                                      dnl, INT32_MAX);
            std::swap(NL, NL2);
        }
        GK_TRY(launch_check("k_frontier"));
        GK_HIP(hipMemcpyAsync(h, dn, 12, hipMemcpyDeviceToHost, ctx->stream));
        GK_HIP(hipStreamSynchronize(ctx->stream));
        nF = h[0];
        nNL = h[2];
        if (full && fixed_k == 0) {
            nl_valid = true;
            if (nNL * 16 < na) nl_window = std::min<int64_t>(nl_window * 2, 1 << 16);
            else if (nNL * 4 > na) nl_window = std::max<int64_t>(nl_window / 2, 2);
        }
        return GK_OK;
    };
    while (alive > 0) {
        if (scan_needed) {
            // A level is finished (or the run starts): every alive edge has
            // support >= k-2.  The next frontier is the alive edges at the
            // minimum support s, at threshold k' - 2 with k' = max(k+1, s+3)
            // (the start uses k itself).  One scan collects the edges below
            // k'-2 for k' = k (+1) and the minimum; another is needed only
            // when that finds nothing.
            const int kt = started ? k + 1 : k;
            bool full = !(nl_valid && kt - 2 < nl_hi);
            GK_TRY(scan(full, kt - 2));
            if (fixed_k > 0 || nF > 0) {
                k = kt;
            } else {
                // an empty near list leaves the minimum to a full scan
                if (!full && nNL == 0) GK_TRY(scan(true, kt - 2));
                const uint32_t mn = ~(uint32_t)h[1];
                k = (int)std::max<int64_t>(kt + 1, (int64_t)mn + 3);
                full = !(nl_valid && k - 2 < nl_hi);
                GK_TRY(scan(full, k - 2));
            }
            started = true;
            scan_needed = false;
        }
        if (nF == 0) {
            if (fixed_k > 0) break;
            scan_needed = true;
            continue;
        }
        // a batch of rounds driven from the device
        hpc->rc[0] = RoundCtl{nP, nF, 0, fc, 0, 0, 0, 0, k, 0, 0, 0, 0};
        hpc->rc[1] = RoundCtl{};
        hpc->rc[1].seq = -1;
        hpc->lvmin = 0;
        hpc->knext = 0;
        hpc->stop = 0;
        hpc->err = 0;
        GK_HIP(hipMemcpyAsync(dpc, hpc, offsetof(PeelCtl, n), hipMemcpyHostToDevice, ctx->stream));
        const PeelDev D{&dpc->stop, &dpc->err, &dpc->lvmin, &dpc->knext, lb, NL, dnl,
                        fixed_k == 0 && nl_valid ? 1 : 0, fixed_k > 0 ? 1u : 0u};
        RoundCtl *dc = dpc->rc;
        for (int r = 0; r < batch; ++r) {
            RoundCtl *ci = dc + (r & 1), *co = dc + ((r + 1) & 1);
            k_round_prep<<<nprep, kPrepThreads, 0, ctx->stream>>>(ci, co, D, ++epoch, fb[0], fb[1], st, tl, eu, ev,
                                                                  sup, arst[cur], fy, fx, flen, fpre, cstart, ccap);
            GK_TRY(stream_begin());
            k_dround_stream<kPeelU><<<gs, 256, 0, ctx->stream>>>(co, D, fb[0], fb[1], fpre, fy, fx, cstart, ccap,
                                                                acol[cur], aeid[cur], hoff, tab, tflt, st, sup,
                                                                &co->nFn,
                                                                near_list(), dctr);
            GK_TRY(stream_end());
        }
        GK_TRY(launch_check("device rounds"));
        GK_HIP(hipMemcpyAsync(hpc, dpc, sizeof(PeelCtl), hipMemcpyDeviceToHost, ctx->stream));
        GK_HIP(hipStreamSynchronize(ctx->stream));
        if (hpc->err) {
            set_error("truss peeling: a round's prefix scan did not complete (look-back wait gave up)");
            return GK_ESTATE;
        }
        RoundCtl rc = hpc->rc[hpc->rc[1].seq > hpc->rc[0].seq ? 1 : 0];
        // slots (rounds and level changes) that ran to the end
        batch = next_batch(batch, hpc->stop ? rc.seq - 1 : batch, 64);
        if (rc.adv) {  // the batch's last slot
            const bool lev = rc.trans && rc.nFn > 0;
            advance_round(rc);
            if (lev) rc.k = hpc->knext;
        }
        k = rc.k;
        if (fixed_k > 0) kbest = k - 1;
        else if (rc.lab > 0) kbest = rc.lab;
        nNL = hpc->nl;
        rounds += rc.rounds;
        alive -= rc.removed;
        // edges marked removed by the batch's preps
        const int64_t marked = (int64_t)nP + rc.removed - rc.nP;
        nP = rc.nP;
        nF = rc.nF;
        fc = rc.cur;
        dead_pending += marked;
        // the alive list is compacted once half of it is gone
        if (alive > 0 && 2 * (alive - nF) < na && !(2 * (dead_pending + nP) * 2 > aal_nz)) {
            GK_HIP(hipMemsetAsync(dac, 0, 4, ctx->stream));
            k_alist_compact<<<grid_scan(ctx, na), 256, 0, ctx->stream>>>(na, alist, st, alist2, dac);
            GK_TRY(launch_check("k_alist_compact"));
            GK_HIP(hipMemcpyAsync(h + 1, dac, 4, hipMemcpyDeviceToHost, ctx->stream));
            GK_HIP(hipStreamSynchronize(ctx->stream));
            na = h[1];
            std::swap(alist, alist2);
        }
        // compact once half of the adjacency entries are dead (a quarter or
        // three quarters measured slower on R-MAT scale 20)
        if (alive > 0 && 2 * (dead_pending + nP) * 2 > aal_nz) {
            k_mark_dead<<<grid_items(ctx, nP), 256, 0, ctx->stream>>>(nP, fb[fc ^ 1], st);
            nP = 0;
            const int nx = cur ^ 1;
            const int gl = (int)std::max<int64_t>(1, std::min<int64_t>((m + kLongThreads - 1) / kLongThreads,
                                                                       (int64_t)ctx->num_cus * 2));
            k_alive_count<<<gw, 256, 0, ctx->stream>>>(m, arst[cur], aeid[cur], st, cnt);
            k_alive_long<false><<<gl, kLongThreads, 0, ctx->stream>>>(m, arst[cur], acol[cur], aeid[cur], st, cnt,
                                                                     nullptr, nullptr, nullptr);
            GK_TRY(exclusive_scan_i32(ctx, cnt, arst[nx], m + 1));
            k_alive_scatter<<<gw, 256, 0, ctx->stream>>>(m, arst[cur], acol[cur], aeid[cur], st, arst[nx], acol[nx],
                                                         aeid[nx]);
            k_alive_long<true><<<gl, kLongThreads, 0, ctx->stream>>>(m, arst[cur], acol[cur], aeid[cur], st, cnt,
                                                                    arst[nx], acol[nx], aeid[nx]);
            GK_HIP(hipMemsetAsync(dac, 0, 4, ctx->stream));
            k_alist_compact<<<grid_scan(ctx, na), 256, 0, ctx->stream>>>(na, alist, st, alist2, dac);
            GK_TRY(launch_check("compaction"));
            GK_HIP(hipMemcpyAsync(h, arst[nx] + m, 4, hipMemcpyDeviceToHost, ctx->stream));
            GK_HIP(hipMemcpyAsync(h + 1, dac, 4, hipMemcpyDeviceToHost, ctx->stream));
            GK_HIP(hipStreamSynchronize(ctx->stream));
            aal_nz = h[0];
            na = h[1];
            std::swap(alist, alist2);
            cur = nx;
            dead_pending = 0;
        }
    }
    // peel-stream figures: device time over all rounds, algorithmic bytes
    // (per streamed entry: column and edge id 8 B, a table probe 8 B, two
    // 2-bit states 0.5 B; per destroyed triangle: two 4-byte support updates)
    unsigned long long hc[2];
    GK_HIP(d2h(ctx, hc, dctr, sizeof hc));
    double sms = 0;
    for (size_t i = 0; i < ev_used; i += 2) {
        float t = 0;
        GK_HIP(hipEventElapsedTime(&t, ev_pool[i], ev_pool[i + 1]));
        sms += t;
    }
    if (ev_used) sms *= (double)stream_launches / (double)(ev_used / 2);
    gk_timing &tt = ctx->last;
    tt.acc_ms[GK_KERNEL_PEEL] += sms;
    tt.acc_bytes[GK_KERNEL_PEEL] += 16.5 * (double)hc[0] + 8.0 * (double)hc[1];
    tt.acc_launches[GK_KERNEL_PEEL] += rounds;
    tt.kernel_ms += sms;
    tt.bytes_alg += 16.5 * (double)hc[0] + 8.0 * (double)hc[1];
    if (rounds_out) *rounds_out = fixed_k > 0 ? rounds + 1 : rounds;
    if (kmax_out) *kmax_out = kbest;
    return GK_OK;
}

}  // namespace
}  // namespace gk

namespace gk {

int edge_index(gk_ctx *ctx, const gk_dcsr *A, EdgeIndex *X) {
    const int64_t m = A->m;
    int32_t *len, *lid;
    GK_TRY(X->alloc(&len, m + 1));
    GK_TRY(X->alloc(&X->lrst, m + 1));
    const int g = (int)std::max<int64_t>(1, std::min<int64_t>((m + 255) / 256, ctx->num_cus * 16));
    k_lower_len<<<g, 256, 0, ctx->stream>>>(m, A->rst, A->col, len);
    GK_TRY(launch_check("k_lower_len"));
    GK_TRY(exclusive_scan_i32(ctx, len, X->lrst, m + 1));
    int32_t nz;
    GK_HIP(d2h(ctx, &nz, X->lrst + m, 4));
    X->m = m;
    X->nL = nz;
    GK_TRY(X->alloc(&X->lcol, nz));
    GK_TRY(X->alloc(&lid, nz));
    GK_TRY(X->alloc(&X->eu, nz));
    GK_TRY(X->alloc(&X->ev, nz));
    GK_TRY(X->alloc(&X->aeid, A->nz));
    const int gw = waves_grid(ctx, m);
    k_lower_copy<<<gw, 256, 0, ctx->stream>>>(m, A->rst, A->col, X->lrst, X->lcol, lid);
    k_edge_ends<<<gw, 256, 0, ctx->stream>>>(m, X->lrst, X->lcol, X->eu, X->ev);
    k_sym_eid<<<gw, 256, 0, ctx->stream>>>(m, A->rst, A->col, X->lrst, X->aeid);
    return launch_check("edge_index");
}

}  // namespace gk

extern "C" int gk_truss_decompose_dev(gk_ctx *ctx, const gk_dcsr *A, uint32_t *d_truss, int32_t *kmax,
                                      const gk_opts *opts) {
    GK_CHECK_ARG(ctx && A && (d_truss || A->nz == 0), "null argument");
    GK_CHECK_ARG(A->m == A->n, "A must be square");
    GK_TRY(require_symmetric(A, "A"));
    gk_opts o = resolve_opts(opts);
    ctx->last = gk_timing{};
    Timer total(ctx);
    total.start();
    TrussBufs B(ctx);
    Lgraph cur;
    GK_TRY(build_lower(ctx, A, B, &cur));
    const int64_t nl = cur.nz;
    uint32_t *tl;
    GK_TRY(B.alloc(&tl, nl));
    int64_t rounds = 0;
    int32_t kb = 0;
    int32_t *sym_eid = nullptr;
    GK_TRY(truss_peel(ctx, A, B, cur, tl, o, &rounds, &kb, 0, &sym_eid));
    if (A->nz) {
        k_gather_u32<<<grid_items(ctx, A->nz), 256, 0, ctx->stream>>>(A->nz, sym_eid, tl, d_truss);
        GK_TRY(launch_check("k_gather_u32"));
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
                    int32_t mid = l + ((h - l) >> 1);
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
    int rc = require_symmetric(&d, "A");
    if (rc) { free_csr(&d); return rc; }
    {
        TrussBufs B(ctx);
        Lgraph cur;
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
            if ((rc = truss_peel(ctx, &d, B, cur, tl, o, &rounds, nullptr, k))) break;
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
    int rc = require_symmetric(&d, "A");
    if (rc) { free_csr(&d); return rc; }
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
