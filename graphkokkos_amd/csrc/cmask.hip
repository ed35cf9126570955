// Complement-mask product C<not M> = A*B (GraphKokkos's mxmm_msa_cmask,
// include/matrix_la.h:733-934): C holds every (i, j) that some product
// term A(i,k)*B(k,j) reaches and M does not hold, with the sum of those
// terms (int32, wrap-around), explicit zeros kept, rows in column order.
//
// Unlike C<M> = A*B the output pattern is not known beforehand, so the
// product runs as expand / sort / reduce:
//   1. k_cm_entries: product terms per A entry (its B row's length);
//      scanned, they give every entry's and every row's first term and cut
//      the rows into chunks of at most kChunkTerms terms;
//   2. per chunk, k_cm_expand writes one (key, value) per term with
//      key = (i - r0) << cb | j (cb = bits of a column, r0 = the chunk's
//      first row); a term at a mask column gets the key (r1 - r0) << cb,
//      above every real key of the chunk.  A radix sort of the pairs over
//      the key's bits only (32-bit keys when they fit) and a reduce-by-key
//      sum the terms per (i, j);
//   3. the chunks' results are concatenated; each row pointer is the
//      position of the row's first key among its chunk's sorted keys.
// The reference's two passes (symbolic row counts, then values through a
// dense B.n accumulator per thread) become sorts here: HBM-bound key
// traffic instead of a B.n-sized scan per row.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce.hpp>
#include <rocprim/device/device_reduce_by_key.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include <vector>

#include "gk_device.h"

using namespace gk;

namespace gk {
namespace {

constexpr int64_t kChunkTerms = int64_t(1) << 27;
constexpr int kMaskLds = 1024;  // mask rows up to this long are searched in LDS
constexpr int kWaves = 4;       // waves per block of the expand kernel
constexpr int64_t kTaskTerms = 4096;  // product terms per expansion task

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Per A entry t: blen[t] = |B row A.col[t]| (0 in a dense row) and arow[t]
// = its row; one wave per row.  blen[A.nz] = 0 closes the scan.
__global__ void k_cm_entries(int64_t m, int64_t nz, const int32_t *__restrict__ arst,
                             const int32_t *__restrict__ acol, const int32_t *__restrict__ brst,
                             const uint8_t *__restrict__ dense, int64_t *__restrict__ blen,
                             int32_t *__restrict__ arow) {
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < m; i += waves)
        for (int32_t p = arst[i] + lane; p < arst[i + 1]; p += 64) {
            const int32_t k = acol[p];
            blen[p] = dense[i] ? 0 : brst[k + 1] - brst[k];
            arow[p] = (int32_t)i;
        }
    if (blockIdx.x == 0 && threadIdx.x == 0) blen[nz] = 0;
}

// foff[i] = aoff[A.rst[i]]: the first product term of row i.
__global__ void k_cm_row_terms(int64_t m, const int32_t *__restrict__ arst, const int64_t *__restrict__ aoff,
                               int64_t *__restrict__ foff) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= m; i += (int64_t)gridDim.x * blockDim.x)
        foff[i] = aoff[arst[i]];
}

// Whether j is among the sorted p[0, len).
__device__ __forceinline__ bool in_sorted(const int32_t *p, int32_t len, int32_t j) {
    int32_t l = 0, h = len;
    while (l < h) {
        const int32_t mid = l + ((h - l) >> 1);
        if (p[mid] < j) l = mid + 1; else h = mid;
    }
    return l < len && p[l] == j;
}

// The chunk's terms [base, base + terms), in A-entry order, cut into tasks
// of kTaskTerms terms, one wave per task: the task's first A entry comes
// from a search over aoff (the scan of B row lengths over A's entries);
// then A entries are taken 64 at a time into LDS (term offset, B row
// start, A value, row, mask row) and every lane takes every 64th term of
// the task in the group, naming its A entry by a search over the offsets.
// Work per wave is bounded whatever the row and B-row lengths.
template <class K>
__global__ __launch_bounds__(64 * kWaves) void k_cm_expand(
    int64_t r0, int64_t base, int64_t terms, int cb, K masked, int64_t anz, const int64_t *__restrict__ aoff,
    const int32_t *__restrict__ arow, const int32_t *__restrict__ acol, const int32_t *__restrict__ aval,
    const int32_t *__restrict__ brst, const int32_t *__restrict__ bcol, const int32_t *__restrict__ bval,
    const int32_t *__restrict__ mrst, const int32_t *__restrict__ mcol, K *__restrict__ key,
    uint32_t *__restrict__ val) {
    __shared__ int64_t s_beg[kWaves][65];  // term offsets of the group's entries (+ the group's end)
    __shared__ int32_t s_b0[kWaves][64];   // B row starts
    __shared__ uint32_t s_a[kWaves][64];   // A values
    __shared__ int32_t s_row[kWaves][64];  // rows
    __shared__ int32_t s_m0[kWaves][64], s_ml[kWaves][64];  // mask rows
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t ntasks = (terms + kTaskTerms - 1) / kTaskTerms;
    for (int64_t task = (int64_t)blockIdx.x * kWaves + w; task < ntasks; task += (int64_t)gridDim.x * kWaves) {
        const int64_t q0 = base + task * kTaskTerms, q1 = min(q0 + kTaskTerms, base + terms);
        // the A entry holding term q0: the last t with aoff[t] <= q0
        int64_t l = 0, h = anz;
        while (l < h) {
            const int64_t mid = l + ((h - l + 1) >> 1);
            if (aoff[mid] <= q0) l = mid; else h = mid - 1;
        }
        for (int64_t t = l;; t += 64) {
            const int64_t te = t + lane;
            wave_sync();
            if (te < anz) {
                const int32_t k = acol[te], i = arow[te];
                s_beg[w][lane] = aoff[te];
                s_b0[w][lane] = brst[k];
                s_a[w][lane] = aval ? (uint32_t)aval[te] : 1u;
                s_row[w][lane] = i;
                s_m0[w][lane] = mrst[i];
                s_ml[w][lane] = mrst[i + 1] - mrst[i];
            } else {
                s_beg[w][lane] = INT64_MAX;
            }
            if (lane == 63) s_beg[w][64] = aoff[min(t + 64, anz)];
            wave_sync();
            const int64_t gend = s_beg[w][64];
            const int64_t lo = max(q0, s_beg[w][0]), hi = min(q1, gend);
            for (int64_t q = lo + lane; q < hi; q += 64) {
                int32_t el = 0, eh = 63;  // the last entry with offset <= q
                while (el < eh) {
                    const int32_t mid = (el + eh + 1) >> 1;
                    if (s_beg[w][mid] <= q) el = mid; else eh = mid - 1;
                }
                const int32_t pb = s_b0[w][el] + (int32_t)(q - s_beg[w][el]);
                const int32_t j = bcol[pb];
                const bool hit = in_sorted(mcol + s_m0[w][el], s_ml[w][el], j);
                key[q - base] = hit ? masked : ((K)(s_row[w][el] - r0) << cb | (K)(uint32_t)j);
                val[q - base] = s_a[w][el] * (bval ? (uint32_t)bval[pb] : 1u);
            }
            if (gend >= q1 || t + 64 >= anz) break;
        }
    }
}

// ---- dense rows ------------------------------------------------------------
// A row with many product terms (relative to its A entries and to C's width)
// is accumulated densely instead: its columns are cut into windows of W
// (a power of two, at most kDenseWindow) columns and one workgroup takes a
// (row, window) pair at a time, with a bitmap of the window's reached
// columns and, in the numeric pass, their int32 sums in LDS.  A first pass
// counts each pair's entries (bitmap only); the second, with the row
// pointers known, writes them in column order straight into C.  Each A
// entry's B row is cut to the window by two binary searches (B rows sorted;
// a single window needs none).
constexpr int kDenseWindowLog = 15;  // 32768 columns: 128 KiB of sums + 4 KiB of bits
constexpr int kDenseThreads = 1024;
// Dense rows are written in one pass into a pool, then copied into C, when
// the pool's upper bound fits this; otherwise a count pass comes first.
constexpr int64_t kPoolMaxBytes = int64_t(8) << 30;
constexpr int kDenseWaves = kDenseThreads / 64;

__global__ void k_cm_flop(int64_t m, const int32_t *__restrict__ arst, const int32_t *__restrict__ acol,
                          const int32_t *__restrict__ brst, int64_t *__restrict__ flop) {
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < m; i += waves) {
        int64_t f = 0;
        for (int32_t p = arst[i] + lane; p < arst[i + 1]; p += 64) f += brst[acol[p] + 1] - brst[acol[p]];
        f = dev::wave_sum(f);
        if (lane == 0) flop[i] = f;
    }
}

// dense[i]: row i takes the dense path.  mode 0: none; 1: every row with
// terms; -1: rows whose terms outweigh the per-window work (a fixed cost
// per window plus two searches per A entry and window).
__global__ void k_cm_classify(int64_t m, int mode, int64_t windows, const int64_t *__restrict__ flop,
                              const int32_t *__restrict__ arst, uint8_t *__restrict__ dense) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t f = flop[i], deg = arst[i + 1] - arst[i];
        bool d = false;
        if (mode == 1) d = f > 0;
        else if (mode < 0) d = f >= windows * (2048 + 16 * deg);
        dense[i] = d ? 1 : 0;
    }
}

__device__ __forceinline__ int32_t lower_bound_col(const int32_t *__restrict__ c, int32_t l, int32_t h, int32_t v) {
    while (l < h) {
        const int32_t mid = l + ((h - l) >> 1);
        if (c[mid] < v) l = mid + 1; else h = mid;
    }
    return l;
}

// What a pass over the (dense row, window) pairs does with a pair's entries.
enum DensePass {
    kDenseCount = 0,  // count them into pcnt
    kDenseWriteC = 1, // write them at rst[row] + poff[pair] (poff: the pair's place in its row)
    kDenseWritePool = 2  // write them at a run of the pool reserved by one atomic on *pool_n;
                         // the run's start into pstart, its length into pcnt
};

template <int PASS>
__global__ __launch_bounds__(kDenseThreads) void k_cm_dense(
    int64_t npairs, int64_t windows, int wlog, int64_t n, const int32_t *__restrict__ drows,
    const int32_t *__restrict__ arst, const int32_t *__restrict__ acol, const int32_t *__restrict__ aval,
    const int32_t *__restrict__ brst, const int32_t *__restrict__ bcol, const int32_t *__restrict__ bval,
    const int32_t *__restrict__ mrst, const int32_t *__restrict__ mcol, int32_t *__restrict__ pcnt,
    const int32_t *__restrict__ poff, const int32_t *__restrict__ rst, int32_t *__restrict__ ccol,
    int32_t *__restrict__ cval, unsigned long long *__restrict__ pool_n, int64_t pool_cap,
    int64_t *__restrict__ pstart, int32_t *__restrict__ over) {
    constexpr bool NUMERIC = PASS != kDenseCount;
    __shared__ uint32_t bits[1 << (kDenseWindowLog - 5)];
    __shared__ uint32_t vals[NUMERIC ? 1 << kDenseWindowLog : 1];
    __shared__ long long s_base;
    __shared__ int64_t s_end[kDenseThreads];  // block scan of the group's B row cuts
    __shared__ int32_t s_b0[kDenseThreads];
    __shared__ uint32_t s_a[kDenseThreads];
    __shared__ int64_t s_red[kDenseWaves];
    __shared__ int32_t s_mlo;
    const int W = 1 << wlog, words = W >> 5;  // wlog <= kDenseWindowLog
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int64_t pair = blockIdx.x; pair < npairs; pair += gridDim.x) {
        const int32_t row = drows[pair / windows];
        const int64_t win = pair % windows;
        const int32_t w0 = (int32_t)(win << wlog);
        const int32_t w1 = (int32_t)min((int64_t)w0 + W, n);
        for (int x = threadIdx.x; x < words; x += kDenseThreads) bits[x] = 0;
        if (NUMERIC)
            for (int x = threadIdx.x; x < W; x += kDenseThreads) vals[x] = 0;
        __syncthreads();
        const int32_t a0 = arst[row], a1 = arst[row + 1];
        if (threadIdx.x == 0) s_mlo = lower_bound_col(mcol, mrst[row], mrst[row + 1], w0);
        // A entries 1024 at a time; their B rows' terms spread over the
        // whole block (a row's hub entries come first and hold most terms)
        for (int32_t g = a0; g < a1; g += kDenseThreads) {
            const int32_t t = g + threadIdx.x;
            int32_t len = 0, s = 0;
            uint32_t a = 0;
            if (t < a1) {
                const int32_t k = acol[t];
                int32_t b0 = brst[k], b1 = brst[k + 1];
                if (w0 > 0) b0 = lower_bound_col(bcol, b0, b1, w0);  // columns are < n
                if (w1 < n) b1 = lower_bound_col(bcol, b0, b1, w1);
                s = b0;
                len = b1 - b0;
                a = aval ? (uint32_t)aval[t] : 1u;
            }
            const int64_t wincl = dev::wave_incl_scan((int64_t)len);
            if (lane == 63) s_red[w] = wincl;
            __syncthreads();
            int64_t before = 0, total = 0;
            for (int x = 0; x < kDenseWaves; ++x) {
                const int64_t r = s_red[x];
                before += x < w ? r : 0;
                total += r;
            }
            s_end[threadIdx.x] = before + wincl;
            s_b0[threadIdx.x] = s;
            s_a[threadIdx.x] = a;
            __syncthreads();
            int32_t el = 0;  // the first entry whose scan exceeds q; q only grows
            for (int64_t q = threadIdx.x; q < total; q += kDenseThreads) {
                if (s_end[el] <= q) {
                    int32_t eh = kDenseThreads - 1;
                    ++el;
                    while (el < eh) {
                        const int32_t mid = (el + eh) >> 1;
                        if (s_end[mid] <= q) el = mid + 1; else eh = mid;
                    }
                }
                const int32_t pb = s_b0[el] + (int32_t)(q - (el ? s_end[el - 1] : 0));
                const int32_t j = bcol[pb] - w0;
                // a B row's consecutive columns share bitmap words and the
                // hub columns recur: test first, so that a set bit costs a
                // (broadcast) read rather than a contended atomic
                const uint32_t bit = 1u << (j & 31);
                if (!(bits[j >> 5] & bit)) atomicOr(&bits[j >> 5], bit);
                if (NUMERIC) atomicAdd(&vals[j], s_a[el] * (bval ? (uint32_t)bval[pb] : 1u));
            }
            __syncthreads();
        }
        __syncthreads();
        {  // the mask row's columns in the window leave the bitmap
            const int32_t m1 = mrst[row + 1];
            for (int32_t p = s_mlo + threadIdx.x; p < m1; p += kDenseThreads) {
                const int32_t j = mcol[p];
                if (j >= w1) break;
                atomicAnd(&bits[(j - w0) >> 5], ~(1u << ((j - w0) & 31)));
            }
        }
        __syncthreads();
        // each thread takes a run of `per` words
        const int per = words > kDenseThreads ? words / kDenseThreads : 1;
        const int wb = threadIdx.x * per;
        int32_t c = 0;
        for (int x = 0; x < per; ++x)
            if (wb + x < words) c += __popc(bits[wb + x]);
        const int32_t incl = dev::wave_incl_scan(c);
        if (lane == 63) s_red[w] = incl;
        __syncthreads();
        if (PASS == kDenseCount) {
            if (threadIdx.x == 0) {
                int64_t tot = 0;
                for (int x = 0; x < kDenseWaves; ++x) tot += s_red[x];
                pcnt[pair] = (int32_t)tot;
            }
        } else {
            int64_t before = 0;
            for (int x = 0; x < w; ++x) before += s_red[x];
            int64_t at;
            if (PASS == kDenseWritePool) {
                if (threadIdx.x == 0) {
                    int64_t tot = 0;
                    for (int x = 0; x < kDenseWaves; ++x) tot += s_red[x];
                    long long b = (long long)atomicAdd(pool_n, (unsigned long long)tot);
                    if (b + tot > pool_cap) {  // cannot happen (the pool holds the bound); never write past it
                        *over = 1;
                        b = -1;
                    }
                    pcnt[pair] = (int32_t)tot;
                    pstart[pair] = b;
                    s_base = b;
                }
                __syncthreads();
                at = s_base;
            } else {
                at = (int64_t)rst[row] + poff[pair];
            }
            int64_t out = at + before + incl - c;
            for (int x = 0; x < per && at >= 0; ++x) {
                if (wb + x >= words) break;
                uint32_t b = bits[wb + x];
                while (b) {
                    const int bit = __ffs(b) - 1;
                    b &= b - 1;
                    const int j = ((wb + x) << 5) + bit;
                    ccol[out] = w0 + j;
                    cval[out] = (int32_t)vals[j];
                    ++out;
                }
            }
        }
        __syncthreads();
    }
}

// Each pair's run of the pool to its place in C, one workgroup per pair.
__global__ void k_cm_dense_place(int64_t npairs, int64_t windows, const int32_t *__restrict__ drows,
                                 const int32_t *__restrict__ pcnt, const int64_t *__restrict__ pstart,
                                 const int32_t *__restrict__ poff, const int32_t *__restrict__ rst,
                                 const int32_t *__restrict__ pcol, const int32_t *__restrict__ pval,
                                 int32_t *__restrict__ ccol, int32_t *__restrict__ cval) {
    for (int64_t pair = blockIdx.x; pair < npairs; pair += gridDim.x) {
        const int64_t src = pstart[pair], dst = (int64_t)rst[drows[pair / windows]] + poff[pair];
        const int32_t len = pcnt[pair];
        if (src < 0) continue;
        for (int32_t x = threadIdx.x; x < len; x += blockDim.x) {
            ccol[dst + x] = pcol[src + x];
            cval[dst + x] = pval[src + x];
        }
    }
}

// Upper bound of the pool: per dense row min(its terms, C's width), as a
// (row, window) pair holds at most min(its terms, the window's width).
__global__ void k_cm_pool_bound(int64_t m, int64_t n, const int64_t *__restrict__ flop,
                                const uint8_t *__restrict__ dense, int64_t *__restrict__ bound) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
        bound[i] = dense[i] ? min(flop[i], n) : 0;
}

// Per dense row: the pairs' places within the row and the row's count.
__global__ void k_cm_dense_rows(int64_t ndense, int64_t windows, const int32_t *__restrict__ drows,
                                const int32_t *__restrict__ pcnt, int32_t *__restrict__ poff,
                                int64_t *__restrict__ cnt) {
    for (int64_t h = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; h < ndense;
         h += (int64_t)gridDim.x * blockDim.x) {
        int64_t run = 0;
        for (int64_t x = 0; x < windows; ++x) {
            poff[h * windows + x] = (int32_t)run;
            run += pcnt[h * windows + x];
        }
        cnt[drows[h]] = run;
    }
}

// ---- sorted rows: counts and placement ----------------------------------
// For the chunk's sorted rows in [r0, r1): cnt[r] = the row's entries among
// the chunk's sorted unique keys and first[r] = the position of its first.
template <class K>
__global__ void k_cm_rowcount(int64_t r0, int64_t r1, int cb, const K *__restrict__ key, int64_t n,
                              const uint8_t *__restrict__ dense, int64_t *__restrict__ cnt,
                              int64_t *__restrict__ first) {
    for (int64_t r = r0 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < r1;
         r += (int64_t)gridDim.x * blockDim.x) {
        if (dense[r]) continue;  // counted by the dense pass
        int64_t b[2];
        for (int e = 0; e < 2; ++e) {
            const K v = (K)(r - r0 + e) << cb;
            int64_t l = 0, h = n;
            while (l < h) {
                const int64_t mid = l + ((h - l) >> 1);
                if (key[mid] < v) l = mid + 1; else h = mid;
            }
            b[e] = l;
        }
        first[r] = b[0];
        cnt[r] = b[1] - b[0];
    }
}

// The chunk's entries to their places in C.
template <class K>
__global__ void k_cm_place(int64_t n, int64_t r0, int cb, const K *__restrict__ key, const uint32_t *__restrict__ v,
                           const int64_t *__restrict__ first, const int64_t *__restrict__ rst64,
                           int32_t *__restrict__ ccol, int32_t *__restrict__ cval) {
    const K cmask = ((K)1 << cb) - 1;
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        const K k = key[p];
        const int64_t r = r0 + (int64_t)(k >> cb);
        const int64_t at = rst64[r] + (p - first[r]);
        ccol[at] = (int32_t)(k & cmask);
        cval[at] = (int32_t)v[p];
    }
}

__global__ void k_cm_rst32(int64_t m, const int64_t *__restrict__ rst64, int32_t *__restrict__ rst) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= m; i += (int64_t)gridDim.x * blockDim.x)
        rst[i] = (int32_t)rst64[i];
}

// Work buffers of one product.  They come from the context's grow-only
// cache (the n-th allocation of a call reuses the n-th buffer, freed by
// gk_finalize), so repeated products allocate nothing once warm; a call
// made while another holds the cache allocates and frees its own.
struct Mem {
    gk_ctx *ctx;
    bool cached;
    size_t next = 0;
    std::vector<void *> own;
    explicit Mem(gk_ctx *c) : ctx(c), cached(!c->cmask_cache_busy) {
        if (cached) ctx->cmask_cache_busy = true;
    }
    ~Mem() {
        if (cached) ctx->cmask_cache_busy = false;
        for (void *x : own) (void)hipFree(x);
    }
    template <class T>
    int alloc(T **out, int64_t n) {
        const size_t bytes = sizeof(T) * (size_t)(n > 0 ? n : 1);
        if (cached) {
            auto &cache = ctx->cmask_cache;
            if (next == cache.size()) cache.emplace_back(nullptr, 0);
            auto &slot = cache[next++];
            if (slot.second < bytes) {
                if (slot.first) {
                    (void)hipStreamSynchronize(ctx->stream);
                    (void)hipFree(slot.first);
                    slot = {nullptr, 0};
                }
                const size_t want = bytes + bytes / 8;  // headroom against regrowth
                if (hipMalloc(&slot.first, want) != hipSuccess) {
                    slot = {nullptr, 0};
                    set_error("hipMalloc of %zu bytes for C<not M> failed", want);
                    return GK_ENOMEM;
                }
                slot.second = want;
            }
            *out = (T *)slot.first;
            return GK_OK;
        }
        if (hipMalloc((void **)out, bytes) != hipSuccess) {
            set_error("hipMalloc of %zu bytes for C<not M> failed", bytes);
            return GK_ENOMEM;
        }
        own.push_back(*out);
        return GK_OK;
    }
};

int grid_of(gk_ctx *ctx, int64_t n, int64_t per_block) {
    return (int)std::max<int64_t>(1, std::min<int64_t>((n + per_block - 1) / per_block, (int64_t)ctx->num_cus * 16));
}

// Bits needed to hold v (v >= 1).
int bits_of(int64_t v) { return 64 - __builtin_clzll((unsigned long long)v); }

// Buffers of the chunk loop, sized for the largest chunk; keys are 32- or
// 64-bit per chunk, so the key buffers hold 8 bytes a term.
struct Work {
    char *key, *skey, *ukey;
    uint32_t *val, *sval, *uval;
    int64_t *nu;
    char *stmp, *rtmp;
    size_t sort_bytes, red_bytes;
};

// One chunk's result: n unique (key, sum) pairs with key = (i - r0) << cb | j.
struct Chunk {
    int64_t r0, r1, n;
    int cb;
    bool wide;  // 64-bit keys
    void *key;
    uint32_t *val;
    double expand_ms, sort_ms;  // device time of the expansion, of sort + reduce
};

template <class K>
int chunk_tmp(gk_ctx *ctx, const Work &W, int64_t terms, size_t *sort_bytes, size_t *red_bytes) {
    GK_HIP(rocprim::radix_sort_pairs(nullptr, *sort_bytes, (K *)W.key, (K *)W.skey, W.val, W.sval, (size_t)terms, 0,
                                     (int)(8 * sizeof(K)), ctx->stream));
    GK_HIP(rocprim::reduce_by_key(nullptr, *red_bytes, (K *)W.skey, W.sval, (size_t)terms, (K *)W.ukey, W.uval, W.nu,
                                  rocprim::plus<uint32_t>(), rocprim::equal_to<K>(), ctx->stream));
    return GK_OK;
}

// Expand, sort over the key's bits, reduce; keeps the unique pairs below
// the mask key in buffers of their own (owned by T).
template <class K>
int run_chunk(gk_ctx *ctx, const gk_dcsr *A, const gk_dcsr *B, const gk_dcsr *M, const int64_t *aoff,
              const int32_t *arow, int64_t base, int64_t terms, Work &W, Mem &T, Chunk &ch) {
    K *key = (K *)W.key, *skey = (K *)W.skey, *ukey = (K *)W.ukey;
    const int end_bit = bits_of(ch.r1 - ch.r0) + ch.cb;
    const int64_t tasks = (terms + kTaskTerms - 1) / kTaskTerms;
    const int blocks = (int)std::min<int64_t>((tasks + kWaves - 1) / kWaves, (int64_t)ctx->num_cus * 16);
    const K masked = (K)(ch.r1 - ch.r0) << ch.cb;
    Timer te(ctx), ts(ctx);
    te.start();
    k_cm_expand<K><<<blocks, 64 * kWaves, 0, ctx->stream>>>(ch.r0, base, terms, ch.cb, masked, A->nz, aoff, arow,
                                                            A->col, A->val, B->rst, B->col, B->val, M->rst, M->col,
                                                            key, W.val);
    GK_TRY(launch_check("k_cm_expand"));
    te.stop();
    ts.start();
    size_t sb = W.sort_bytes, rb = W.red_bytes;
    GK_HIP(rocprim::radix_sort_pairs(W.stmp, sb, key, skey, W.val, W.sval, (size_t)terms, 0, end_bit, ctx->stream));
    GK_HIP(rocprim::reduce_by_key(W.rtmp, rb, skey, W.sval, (size_t)terms, ukey, W.uval, W.nu,
                                  rocprim::plus<uint32_t>(), rocprim::equal_to<K>(), ctx->stream));
    ts.stop();
    int64_t hn = 0;
    GK_HIP(d2h(ctx, &hn, W.nu, sizeof hn));
    K last = 0;
    GK_HIP(d2h(ctx, &last, ukey + hn - 1, sizeof last));
    if (last == masked) --hn;  // the masked terms' key
    ch.n = hn;
    ch.expand_ms = te.ms();
    ch.sort_ms = ts.ms();
    if (hn == 0) return GK_OK;
    K *k2;
    uint32_t *v2;
    GK_TRY(T.alloc(&k2, hn));
    GK_TRY(T.alloc(&v2, hn));
    GK_HIP(hipMemcpyAsync(k2, ukey, sizeof(K) * (size_t)hn, hipMemcpyDeviceToDevice, ctx->stream));
    GK_HIP(hipMemcpyAsync(v2, W.uval, 4 * (size_t)hn, hipMemcpyDeviceToDevice, ctx->stream));
    ch.key = k2;
    ch.val = v2;
    return GK_OK;
}

}  // namespace
}  // namespace gk

extern "C" int gk_masked_spgemm_cmask_dev(gk_ctx *ctx, const gk_dcsr *A, const gk_dcsr *B, const gk_dcsr *M,
                                          int32_t dense_rows, gk_dcsr **C) {
    GK_CHECK_ARG(ctx && A && B && M && C, "null argument");
    GK_CHECK_ARG(*C == nullptr || (*C != A && *C != B && *C != M && (*C)->cap_rst > 0),
                 "*C must be NULL or a result of an earlier gk_masked_spgemm_cmask_dev (not an operand)");
    GK_CHECK_ARG(dense_rows >= -1 && dense_rows <= 2, "dense_rows must be -1, 0, 1 or 2");
    if (A->n != B->m || A->m != M->m || B->n != M->n) {
        set_error("shape mismatch: A %lldx%lld, B %lldx%lld, M %lldx%lld", (long long)A->m, (long long)A->n,
                  (long long)B->m, (long long)B->n, (long long)M->m, (long long)M->n);
        return GK_EINVAL;
    }
    GK_TRY(require_sorted(M, "M"));
    const int64_t m = A->m, n = M->n, anz = A->nz;
    Timer total(ctx), tdense(ctx);
    total.start();
    Mem T(ctx);
    // dense windows: W = the power of two >= n, at most 2^kDenseWindowLog
    int wlog = 5;
    while (wlog < kDenseWindowLog && ((int64_t)1 << wlog) < n) ++wlog;
    const int64_t windows = std::max<int64_t>(1, (n + ((int64_t)1 << wlog) - 1) >> wlog);
    // a window cut of B rows needs them sorted
    const int mode = (windows > 1 && !B->sorted) ? 0 : dense_rows == 2 ? 1 : dense_rows;

    // 1. terms per row, the dense rows
    int64_t *flop;
    uint8_t *dflag;
    int32_t *drows, *nd32;
    GK_TRY(T.alloc(&flop, m + 1));
    GK_TRY(T.alloc(&dflag, m + 1));
    GK_TRY(T.alloc(&drows, m + 1));
    GK_TRY(T.alloc(&nd32, 1));
    k_cm_flop<<<grid_of(ctx, m, 4), 256, 0, ctx->stream>>>(m, A->rst, A->col, B->rst, flop);
    k_cm_classify<<<grid_of(ctx, m, 256), 256, 0, ctx->stream>>>(m, mode, windows, flop, A->rst, dflag);
    GK_TRY(launch_check("k_cm_classify"));
    int64_t ndense = 0, terms_all = 0;
    if (m > 0) {  // all product terms, for the timing record
        int64_t *tsum;
        size_t tmp = 0;
        GK_TRY(T.alloc(&tsum, 1));
        GK_HIP(rocprim::reduce(nullptr, tmp, flop, tsum, (int64_t)0, (size_t)m, rocprim::plus<int64_t>(),
                               ctx->stream));
        char *t;
        GK_TRY(T.alloc(&t, (int64_t)tmp));
        GK_HIP(rocprim::reduce(t, tmp, flop, tsum, (int64_t)0, (size_t)m, rocprim::plus<int64_t>(), ctx->stream));
        GK_HIP(d2h(ctx, &terms_all, tsum, sizeof terms_all));
    }
    if (mode != 0 && m > 0) {
        size_t tmp = 0;
        rocprim::counting_iterator<int32_t> ids(0);
        GK_HIP(rocprim::select(nullptr, tmp, ids, dflag, drows, nd32, (size_t)m, ctx->stream));
        char *t;
        GK_TRY(T.alloc(&t, (int64_t)tmp));
        GK_HIP(rocprim::select(t, tmp, ids, dflag, drows, nd32, (size_t)m, ctx->stream));
        int32_t h = 0;
        GK_HIP(d2h(ctx, &h, nd32, sizeof h));
        ndense = h;
    }
    // the dense rows' entries are bounded by min(terms, n) per row: with a
    // pool of that size they take one pass (written there, then copied)
    int64_t pool_bound = 0;
    if (ndense > 0) {
        int64_t *bnd, *bsum;
        GK_TRY(T.alloc(&bnd, m));
        GK_TRY(T.alloc(&bsum, 1));
        k_cm_pool_bound<<<grid_of(ctx, m, 256), 256, 0, ctx->stream>>>(m, n, flop, dflag, bnd);
        GK_TRY(launch_check("k_cm_pool_bound"));
        size_t tmp = 0;
        GK_HIP(rocprim::reduce(nullptr, tmp, bnd, bsum, (int64_t)0, (size_t)m, rocprim::plus<int64_t>(),
                               ctx->stream));
        char *t;
        GK_TRY(T.alloc(&t, (int64_t)tmp));
        GK_HIP(rocprim::reduce(t, tmp, bnd, bsum, (int64_t)0, (size_t)m, rocprim::plus<int64_t>(), ctx->stream));
        GK_HIP(d2h(ctx, &pool_bound, bsum, sizeof pool_bound));
    }
    const bool one_pass = ndense > 0 && dense_rows != 2 && pool_bound * 8 <= kPoolMaxBytes;

    // 2. the sorted rows' terms: per A entry (0 in dense rows), scanned
    int64_t *blen, *aoff, *foff;
    int32_t *arow;
    GK_TRY(T.alloc(&blen, anz + 1));
    GK_TRY(T.alloc(&aoff, anz + 1));
    GK_TRY(T.alloc(&arow, anz));
    GK_TRY(T.alloc(&foff, m + 1));
    k_cm_entries<<<grid_of(ctx, m, 4), 256, 0, ctx->stream>>>(m, anz, A->rst, A->col, B->rst, dflag, blen, arow);
    GK_TRY(launch_check("k_cm_entries"));
    {
        size_t tmp = 0;
        GK_HIP(rocprim::exclusive_scan(nullptr, tmp, blen, aoff, (int64_t)0, (size_t)(anz + 1),
                                       rocprim::plus<int64_t>(), ctx->stream));
        char *t;
        GK_TRY(T.alloc(&t, (int64_t)tmp));
        GK_HIP(rocprim::exclusive_scan(t, tmp, blen, aoff, (int64_t)0, (size_t)(anz + 1), rocprim::plus<int64_t>(),
                                       ctx->stream));
    }
    k_cm_row_terms<<<grid_of(ctx, m + 1, 256), 256, 0, ctx->stream>>>(m, A->rst, aoff, foff);
    GK_TRY(launch_check("k_cm_row_terms"));
    std::vector<int64_t> hoff(m + 1);
    GK_HIP(d2h(ctx, hoff.data(), foff, sizeof(int64_t) * (size_t)(m + 1)));

    // 3. the dense rows: their entries into the pool (or their counts
    // only), while the host cuts chunks
    int32_t *pcnt, *poff, *pcol = nullptr, *pval = nullptr, *over = nullptr;
    int64_t *cnt, *first, *pstart = nullptr;
    unsigned long long *pool_n = nullptr;
    const int64_t npairs = ndense * windows;
    GK_TRY(T.alloc(&pcnt, npairs));
    GK_TRY(T.alloc(&poff, npairs));
    GK_TRY(T.alloc(&cnt, m + 1));
    GK_TRY(T.alloc(&first, m + 1));
    GK_HIP(hipMemsetAsync(cnt, 0, sizeof(int64_t) * (size_t)(m + 1), ctx->stream));
    if (one_pass) {
        GK_TRY(T.alloc(&pcol, pool_bound));
        GK_TRY(T.alloc(&pval, pool_bound));
        GK_TRY(T.alloc(&pstart, npairs));
        GK_TRY(T.alloc(&pool_n, 1));
        GK_TRY(T.alloc(&over, 1));
        GK_HIP(hipMemsetAsync(pool_n, 0, sizeof *pool_n, ctx->stream));
        GK_HIP(hipMemsetAsync(over, 0, sizeof *over, ctx->stream));
    }
    tdense.start();
    if (npairs && one_pass) {
        const int g = (int)std::min<int64_t>(npairs, (int64_t)ctx->num_cus);
        k_cm_dense<kDenseWritePool><<<g, kDenseThreads, 0, ctx->stream>>>(
            npairs, windows, wlog, n, drows, A->rst, A->col, A->val, B->rst, B->col, B->val, M->rst, M->col, pcnt,
            nullptr, nullptr, pcol, pval, pool_n, pool_bound, pstart, over);
        GK_TRY(launch_check("k_cm_dense<pool>"));
    } else if (npairs) {
        const int g = (int)std::min<int64_t>(npairs, (int64_t)ctx->num_cus * 2);
        k_cm_dense<kDenseCount><<<g, kDenseThreads, 0, ctx->stream>>>(
            npairs, windows, wlog, n, drows, A->rst, A->col, A->val, B->rst, B->col, B->val, M->rst, M->col, pcnt,
            nullptr, nullptr, nullptr, nullptr, nullptr, 0, nullptr, nullptr);
        GK_TRY(launch_check("k_cm_dense<count>"));
    }
    if (npairs) {
        k_cm_dense_rows<<<grid_of(ctx, ndense, 256), 256, 0, ctx->stream>>>(ndense, windows, drows, pcnt, poff, cnt);
        GK_TRY(launch_check("k_cm_dense_rows"));
    }
    tdense.stop();
    double dense_ms = 0;

    // row chunks of at most kChunkTerms sorted-row terms (a longer row is a chunk alone)
    std::vector<int64_t> cuts{0};
    while (cuts.back() < m) {
        const int64_t r0 = cuts.back();
        int64_t r1 = std::upper_bound(hoff.begin() + r0 + 1, hoff.end(), hoff[r0] + kChunkTerms) - hoff.begin() - 1;
        if (r1 <= r0) r1 = r0 + 1;
        cuts.push_back(std::min(r1, m));
    }
    int64_t maxterms = 0;
    for (size_t c = 1; c < cuts.size(); ++c) maxterms = std::max(maxterms, hoff[cuts[c]] - hoff[cuts[c - 1]]);
    if (maxterms > INT32_MAX) {
        set_error("a row of C<not M> has %lld product terms (int32 limit)", (long long)maxterms);
        return GK_ERANGE;
    }
    // values as uint32: the sums wrap around like the reference's int arithmetic
    const int cb = bits_of(std::max<int64_t>(n - 1, 1));
    Work W{};
    GK_TRY(T.alloc(&W.key, 8 * maxterms));
    GK_TRY(T.alloc(&W.skey, 8 * maxterms));
    GK_TRY(T.alloc(&W.ukey, 8 * maxterms));
    GK_TRY(T.alloc(&W.val, maxterms));
    GK_TRY(T.alloc(&W.sval, maxterms));
    GK_TRY(T.alloc(&W.uval, maxterms));
    GK_TRY(T.alloc(&W.nu, 1));
    {
        size_t s32 = 0, s64 = 0, r32 = 0, r64 = 0;
        GK_TRY(chunk_tmp<uint32_t>(ctx, W, maxterms, &s32, &r32));
        GK_TRY(chunk_tmp<uint64_t>(ctx, W, maxterms, &s64, &r64));
        W.sort_bytes = std::max(s32, s64);
        W.red_bytes = std::max(r32, r64);
        GK_TRY(T.alloc(&W.stmp, (int64_t)W.sort_bytes));
        GK_TRY(T.alloc(&W.rtmp, (int64_t)W.red_bytes));
    }
    // 4. sorted rows: per chunk expand / sort / reduce, then the rows' counts
    std::vector<Chunk> chunks;
    double expand_ms = 0, sort_ms = 0, bytes = 0;
    int64_t runs = 0;
    for (size_t c = 1; c < cuts.size(); ++c) {
        Chunk ch{};
        ch.r0 = cuts[c - 1];
        ch.r1 = cuts[c];
        ch.cb = cb;
        const int64_t terms = hoff[ch.r1] - hoff[ch.r0];
        if (terms == 0) continue;  // no sorted-row entries: counts stay 0
        ch.wide = bits_of(ch.r1 - ch.r0) + cb > 32;
        GK_TRY(ch.wide ? run_chunk<uint64_t>(ctx, A, B, M, aoff, arow, hoff[ch.r0], terms, W, T, ch)
                       : run_chunk<uint32_t>(ctx, A, B, M, aoff, arow, hoff[ch.r0], terms, W, T, ch));
        ++runs;
        expand_ms += ch.expand_ms;
        sort_ms += ch.sort_ms;
        // expansion bytes: per term B's column (and value) read, key and
        // value written
        bytes += (double)terms * (4 + (B->val ? 4 : 0) + (ch.wide ? 8 : 4) + 4);
        if (ch.n == 0) continue;
        const int gr = grid_of(ctx, ch.r1 - ch.r0, 256);
        if (ch.wide)
            k_cm_rowcount<uint64_t><<<gr, 256, 0, ctx->stream>>>(ch.r0, ch.r1, cb, (const uint64_t *)ch.key, ch.n,
                                                                 dflag, cnt, first);
        else
            k_cm_rowcount<uint32_t><<<gr, 256, 0, ctx->stream>>>(ch.r0, ch.r1, cb, (const uint32_t *)ch.key, ch.n,
                                                                 dflag, cnt, first);
        GK_TRY(launch_check("k_cm_rowcount"));
        chunks.push_back(ch);
    }
    dense_ms += tdense.ms();

    // 5. row pointers
    int64_t *rst64;
    GK_TRY(T.alloc(&rst64, m + 1));
    {
        size_t tmp = 0;
        GK_HIP(rocprim::exclusive_scan(nullptr, tmp, cnt, rst64, (int64_t)0, (size_t)(m + 1),
                                       rocprim::plus<int64_t>(), ctx->stream));
        char *t;
        GK_TRY(T.alloc(&t, (int64_t)tmp));
        GK_HIP(rocprim::exclusive_scan(t, tmp, cnt, rst64, (int64_t)0, (size_t)(m + 1), rocprim::plus<int64_t>(),
                                       ctx->stream));
    }
    int64_t nz = 0;
    GK_HIP(d2h(ctx, &nz, rst64 + m, sizeof nz));
    if (nz > INT32_MAX) {
        set_error("C<not M> has %lld entries (int32 CSR limit)", (long long)nz);
        return GK_ERANGE;
    }
    // the result: a new CSR, or the caller's from an earlier call refilled,
    // its storage grown only when too small (as the reference's resizeVals)
    gk_dcsr *d = *C;
    const bool fresh = d == nullptr;
    if (fresh) d = new gk_dcsr();
    auto fail = [&](int rc) {
        if (fresh) {
            free_csr(d);
            delete d;
        } else {
            d->m = d->n = d->nz = 0;  // left empty but valid for gk_dcsr_free
        }
        return rc;
    };
    d->ctx = ctx;
    d->owns = true;
    d->sorted = true;
    d->lower = d->symmetric = false;
    if (d->cap_rst < m + 1) {
        if (d->rst) (void)hipFree(d->rst);
        d->rst = nullptr;
        d->cap_rst = 0;
        if (hipMalloc(&d->rst, 4 * (size_t)(m + 1)) != hipSuccess) {
            d->rst = nullptr;
            set_error("hipMalloc of C<not M> failed");
            return fail(GK_ENOMEM);
        }
        d->cap_rst = m + 1;
    }
    const int64_t want_nz = std::max<int64_t>(nz, 1);
    if (d->cap_nz < want_nz || !d->col || !d->val) {
        if (d->col) (void)hipFree(d->col);
        if (d->val) (void)hipFree(d->val);
        d->col = d->val = nullptr;
        d->cap_nz = 0;
        if (hipMalloc(&d->col, 4 * (size_t)want_nz) != hipSuccess ||
            hipMalloc(&d->val, 4 * (size_t)want_nz) != hipSuccess) {
            set_error("hipMalloc of C<not M> failed");
            return fail(GK_ENOMEM);
        }
        d->cap_nz = want_nz;
    }
    d->m = m;
    d->n = n;
    d->nz = nz;
    k_cm_rst32<<<grid_of(ctx, m + 1, 256), 256, 0, ctx->stream>>>(m, rst64, d->rst);
    // 6. the entries to their places: sorted rows, then dense rows (second pass)
    for (const Chunk &ch : chunks) {
        const int g = grid_of(ctx, ch.n, 256);
        if (ch.wide)
            k_cm_place<uint64_t><<<g, 256, 0, ctx->stream>>>(ch.n, ch.r0, cb, (const uint64_t *)ch.key, ch.val, first,
                                                              rst64, d->col, d->val);
        else
            k_cm_place<uint32_t><<<g, 256, 0, ctx->stream>>>(ch.n, ch.r0, cb, (const uint32_t *)ch.key, ch.val, first,
                                                              rst64, d->col, d->val);
    }
    int rc;
    if ((rc = launch_check("k_cm_place")) != GK_OK) return fail(rc);
    if (npairs) {
        tdense.start();
        if (one_pass) {
            int32_t o = 0;
            if (d2h(ctx, &o, over, sizeof o) != hipSuccess) return fail(GK_EHIP);
            if (o) {
                set_error("C<not M>: dense rows exceeded their pool bound (internal error)");
                return fail(GK_ESTATE);
            }
            const int g = (int)std::min<int64_t>(npairs, (int64_t)ctx->num_cus * 8);
            k_cm_dense_place<<<g, 256, 0, ctx->stream>>>(npairs, windows, drows, pcnt, pstart, poff, d->rst, pcol,
                                                         pval, d->col, d->val);
            if ((rc = launch_check("k_cm_dense_place")) != GK_OK) return fail(rc);
        } else {
            const int g = (int)std::min<int64_t>(npairs, (int64_t)ctx->num_cus);
            k_cm_dense<kDenseWriteC><<<g, kDenseThreads, 0, ctx->stream>>>(
                npairs, windows, wlog, n, drows, A->rst, A->col, A->val, B->rst, B->col, B->val, M->rst, M->col,
                nullptr, poff, d->rst, d->col, d->val, nullptr, 0, nullptr, nullptr);
            if ((rc = launch_check("k_cm_dense<write>")) != GK_OK) return fail(rc);
        }
        tdense.stop();
    }
    total.stop();
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
        set_error("C<not M> failed on the device");
        return fail(GK_EHIP);
    }
    if (npairs) dense_ms += tdense.ms();
    gk_timing &tm = ctx->last;
    tm = gk_timing{};
    tm.total_ms = total.ms();
    tm.launches = runs;
    tm.iterations = ndense;
    tm.flop = (double)terms_all;
    tm.hits = (double)nz;
    tm.kernel_ms = expand_ms;
    tm.select_ms = sort_ms;
    tm.acc_ms[0] = dense_ms;
    tm.acc_launches[0] = npairs ? 2 : 0;
    {
        // dense rows: per term B's column (and value) read, in the count
        // pass too when there is one; per entry its column and value
        // written once, or to the pool and then copied (read and written
        // again); the A entries and mask rows, a few per cent, left out
        int64_t sorted_nz = 0;
        for (const Chunk &ch : chunks) sorted_nz += ch.n;
        const double dterms = (double)(terms_all - hoff[m]), dnz = (double)(nz - sorted_nz);
        tm.acc_bytes[0] = one_pass ? dterms * (4 + (B->val ? 4 : 0)) + dnz * 24
                                   : dterms * (4 + 4 + (B->val ? 4 : 0)) + dnz * 8;
    }
    // plus, per A entry, its column (and value) and B's two row pointers;
    // the mask rows and the row offsets once
    tm.bytes_alg = bytes + (double)A->nz * (4 + (A->val ? 4 : 0) + 8) + (double)M->nz * 4 + 16.0 * (double)m;
    *C = d;
    return GK_OK;
}

extern "C" int gk_dcsr_download_val(gk_ctx *ctx, const gk_dcsr *d, int32_t *val) {
    GK_CHECK_ARG(ctx && d && (val || d->nz == 0), "null argument");
    GK_CHECK_ARG(d->val || d->nz == 0, "the CSR holds a pattern (no values)");
    if (d->nz) GK_HIP(d2h(ctx, val, d->val, sizeof(int32_t) * (size_t)d->nz));
    return GK_OK;
}
