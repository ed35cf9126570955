// Masked sparse-sparse product C<M> = A*B on gfx950, and the triangle-count
// and edge-support products built on it.
//
// GraphKokkos computes the same product with a dense per-thread accumulator
// of B.n entries (MSA, include/matrix_la.h:576-635) or a mask-indexed
// accumulator with a merge walk (MCA, include/matrix_la.h:493-528).  Here:
//
//  1. k_row_stats: a tile of 8 lanes per row records mask length, A-row length,
//     flop = sum_k |B_k| and max |B_k|.
//  2. Every row gets an accumulator (fixed thresholds, the learned cost
//     model, or a forced choice) and is cut into tasks of at most
//     kMaxTaskA A-entries and about opts.split_flop streamed B entries.
//  3. Tasks are counting-sorted into bins (accumulator, log2 flop
//     descending) so heavy tasks start first.
//  4. One persistent kernel per accumulator drains its bins:
//       MSA-LDS      the mask row hashed into LDS (key -> mask position),
//                    dense per-position counters in LDS;
//       MCA-wave     the sorted mask row in LDS, each lane binary-searches
//                    its streamed B entry, hits counted by ballot/popcount;
//                    16-lane tiles, four mask rows per wave;
//       hash-global  the same table as MSA-LDS but in HBM scratch, for mask
//                    rows too long for LDS;
//       dense-global a column-indexed map of the mask row in HBM scratch
//                    (B.n entries per wave), for the longest mask rows.
//     Each task streams the concatenation of its B rows with all 64 lanes:
//     lane g takes streamed entries g, g+64, ..., locating its B row in an
//     LDS prefix of B-row lengths, so reads of B are coalesced whatever the
//     row lengths.
//
// Modes: TC counts hits (C<L>=L*L, sum of C; src/main.cpp:69-90), VALUES
// writes C's values at M's pattern (mxmm_msa / mxmm_mca semantics), SUPPORT
// turns every hit (i,k,j) of C<L>=L*L into +1 on the three edges of the
// triangle (edge supports of the symmetric graph, S<A>=A*A restricted to L).
#include "gk_device.h"
#include "spgemm.h"

using namespace gk;

namespace gk {

// ---------------------------------------------------------------------------
// features and task construction

// Tiles of G lanes per row (rows average a few dozen entries): each lane
// walks every G-th A entry, then the tile reduces.
template <int G, class T>
__device__ __forceinline__ T tile_sum(T v) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
    return v;
}

template <int G>
__global__ __launch_bounds__(256) void k_row_stats(int64_t m, const int32_t *__restrict__ arst,
                                                   const int32_t *__restrict__ acol,
                                                   const int32_t *__restrict__ brst,
                                                   const int32_t *__restrict__ mrst, RowStats st,
                                                   int32_t hub_n, int same_am, int32_t probe) {
    const int gl = threadIdx.x % G;
    const int64_t per_block = blockDim.x / G;
    for (int64_t i = (int64_t)blockIdx.x * per_block + threadIdx.x / G; i < m; i += (int64_t)gridDim.x * per_block) {
        const int32_t a0 = arst[i], a1 = arst[i + 1];
        int64_t flop = 0, sflop = 0, mflop = 0;
        int32_t bmax = 0;
        for (int32_t p = a0 + gl; p < a1; p += G) {
            const int32_t k = acol[p];
            const int32_t len = brst[k + 1] - brst[k];
            sflop += len;
            // window kernels: the reads of hub_choice's way (bit tests,
            // bitmap words or entries); with A the mask, the mask entries
            // below k are the p - a0 before it
            const int32_t words = (k + 31) >> 5, below = p - a0;
            if (k < hub_n && len > 0 && same_am && probe > 0 && probe * below < len && probe * below < words)
                flop += below;
            else
                flop += hub_row(hub_n, k, len) ? words : len;
            mflop += same_am && k < hub_n ? min(len, below) : len;
            bmax = max(bmax, len);
        }
        flop = tile_sum<G>(flop);
        sflop = tile_sum<G>(sflop);
        mflop = tile_sum<G>(mflop);
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) bmax = max(bmax, __shfl_xor(bmax, o, G));
        if (gl == 0) {
            st.mlen[i] = mrst[i + 1] - mrst[i];
            st.alen[i] = a1 - a0;
            st.flop[i] = flop;
            st.sflop[i] = sflop;
            st.mflop[i] = mflop;
            st.bmax[i] = bmax;
        }
    }
}

// Capacity rules shared by the host and the kernels.
__host__ __device__ __forceinline__ int lds_fits(int32_t mlen) { return mlen <= kLdsMaxMask; }
__host__ __device__ __forceinline__ int mca_fits(int32_t mlen) { return mlen <= kMcaMaxMask; }

__device__ __forceinline__ int flop_bin(int64_t flop) {
    int lg = flop > 0 ? 63 - __clzll((unsigned long long)flop) : 0;
    return lg > kFlopBins - 1 ? kFlopBins - 1 : lg;
}

// Work estimate of row i in kernel class cls: the window kernels read hub
// rows as bitmap words, MCA-wave tests single bits of them, the global
// kernels stream every B entry.
__device__ __forceinline__ int64_t class_flop(const SelectParams &sp, int cls, const RowStats &st, int64_t i) {
    if (!sp.hub) return st.sflop[i];
    if (cls == GK_ACC_MSA_LDS || cls == kClassMsaLarge) return st.flop[i];
    if (cls == GK_ACC_MCA_WAVE) return st.mflop[i];
    return st.sflop[i];
}

__device__ __forceinline__ int32_t task_count(const SelectParams &sp, int cls, int32_t alen, int64_t flop) {
    const int32_t maxa = cls == GK_ACC_MCA_WAVE ? kMcaMaxTaskA : kMaxTaskA;
    int64_t by_a = (alen + maxa - 1) / maxa;
    int64_t by_f = (flop + sp.split_flop - 1) / sp.split_flop;
    int64_t t = by_a > by_f ? by_a : by_f;
    if (t > alen) t = alen;
    return (int32_t)(t < 1 ? 1 : t);
}

// Pick the accumulator for every row (or take the learned choice already in
// st.acc), count its tasks and add them to the bin histogram and the
// per-class work counters.  Rows of a wave usually fall into one or two bins:
// the wave reduces each distinct bin with ballots and shuffles, and only
// that bin's leader lane touches the block's LDS counters.
__global__ __launch_bounds__(256) void k_select_bin(int64_t m, RowStats st, SelectParams sp,
                                                    int32_t *__restrict__ bin_count,
                                                    int32_t *__restrict__ acc_maxmask,
                                                    unsigned long long *__restrict__ stats) {
    __shared__ int32_t h[kNumBins];
    __shared__ int32_t mx[kNumClass];
    __shared__ unsigned long long ss[kNumClass * ST_NUM];
    for (int b = threadIdx.x; b < kNumBins; b += blockDim.x) h[b] = 0;
    for (int b = threadIdx.x; b < kNumClass * ST_NUM; b += blockDim.x) ss[b] = 0;
    for (int b = threadIdx.x; b < kNumClass; b += blockDim.x) mx[b] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    for (int64_t base = blockIdx.x * (int64_t)blockDim.x; base < m; base += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = base + threadIdx.x;
        int acc = -1, bin = 0;
        int32_t mlen = 0, alen = 0, nt = 0;
        int64_t flop = 0;
        if (i < m) {
            mlen = st.mlen[i];
            flop = st.sflop[i];
            alen = st.alen[i];
            if (mlen > 0 && flop > 0) {
                if (sp.selector == GK_SEL_FORCED) acc = sp.forced;
                else if (sp.selector == GK_SEL_LEARNED || sp.selector == kSelMixed) acc = st.acc[i];
                else acc = heuristic_accum(sp, mlen);
                // Capacity overrides: a choice that cannot hold the mask row
                // falls to the next accumulator that can.
                if (acc == GK_ACC_MCA_WAVE && !mca_fits(mlen)) acc = GK_ACC_MSA_LDS;
                if (acc == GK_ACC_MSA_LDS && !lds_fits(mlen)) acc = GK_ACC_HASH_GLOBAL;
                if (acc == GK_ACC_HASH_GLOBAL && mlen > kHashGlobalMaxMask) acc = GK_ACC_DENSE_GLOBAL;
                if (acc == GK_ACC_MSA_LDS && mlen > sp.small_max_mask) acc = kClassMsaLarge;
                flop = class_flop(sp, acc, st, i);
                nt = task_count(sp, acc, alen, flop);
                bin = acc * kFlopBins + (kFlopBins - 1 - flop_bin(flop));
            }
            st.acc[i] = (int8_t)acc;  // kernel class from here on
            st.ntask[i] = nt;
        }
        bool pend = acc >= 0;
        for (;;) {
            const uint64_t act = __ballot(pend);
            if (!act) break;
            const int leader = __ffsll((unsigned long long)act) - 1;
            const int lbin = __shfl(bin, leader, 64);
            const bool mine = pend && bin == lbin;
            const int32_t s_nt = dev::wave_sum(mine ? nt : 0);
            const unsigned long long s_m = dev::wave_sum((unsigned long long)(mine ? mlen : 0));
            const unsigned long long s_a = dev::wave_sum((unsigned long long)(mine ? alen : 0));
            const unsigned long long s_f = dev::wave_sum((unsigned long long)(mine ? flop : 0));
            const int32_t s_x = dev::wave_max(mine ? mlen : 0);
            if (lane == leader) {
                const int c = lbin / kFlopBins;
                atomicAdd(&h[lbin], s_nt);
                atomicMax(&mx[c], s_x);
                atomicAdd(&ss[c * ST_NUM + ST_TASKS], (unsigned long long)s_nt);
                atomicAdd(&ss[c * ST_NUM + ST_MASK], s_m);
                atomicAdd(&ss[c * ST_NUM + ST_AROW], s_a);
                atomicAdd(&ss[c * ST_NUM + ST_FLOP], s_f);
            }
            pend = pend && !mine;
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < kNumBins; b += blockDim.x)
        if (h[b]) atomicAdd(&bin_count[b], h[b]);
    for (int b = threadIdx.x; b < kNumClass; b += blockDim.x)
        if (mx[b]) atomicMax(&acc_maxmask[b], mx[b]);
    for (int b = threadIdx.x; b < kNumClass * ST_NUM; b += blockDim.x)
        if (ss[b]) atomicAdd(&stats[b], ss[b]);
}

// Exclusive scan of the bin histogram (one block), also per-accumulator
// ranges and per-accumulator maximum mask length.
__global__ void k_bin_scan(const int32_t *__restrict__ bin_count, int32_t *__restrict__ bin_off,
                           int32_t *__restrict__ bin_cur) {
    if (threadIdx.x == 0) {
        int32_t s = 0;
        for (int b = 0; b < kNumBins; ++b) {
            bin_off[b] = s;
            bin_cur[b] = s;
            s += bin_count[b];
        }
        bin_off[kNumBins] = s;
    }
}

// Tasks of every row written into their bins.  A block first counts its
// rows' tasks per bin in LDS, reserves each bin's span with one global atomic,
// then writes -- so the ~100 bin cursors see one atomic per block, not per row.
__global__ __launch_bounds__(256) void k_scatter_tasks(int64_t m, RowStats st, SelectParams sp,
                                                       const int32_t *__restrict__ arst,
                                                       int32_t *__restrict__ bin_cur, Task *__restrict__ tasks) {
    __shared__ int32_t lcnt[kNumBins];
    __shared__ int32_t lbase[kNumBins];
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < m; base += (int64_t)gridDim.x * blockDim.x) {
        for (int b = threadIdx.x; b < kNumBins; b += blockDim.x) lcnt[b] = 0;
        __syncthreads();
        const int64_t i = base + threadIdx.x;
        int cls = -1, bin = 0;
        int32_t nt = 0, off = 0;
        if (i < m) {
            cls = st.acc[i];
            if (cls >= 0) {
                nt = st.ntask[i];
                bin = cls * kFlopBins + (kFlopBins - 1 - flop_bin(class_flop(sp, cls, st, i)));
                off = atomicAdd(&lcnt[bin], nt);
            }
        }
        __syncthreads();
        for (int b = threadIdx.x; b < kNumBins; b += blockDim.x)
            if (lcnt[b]) lbase[b] = atomicAdd(&bin_cur[b], lcnt[b]);
        __syncthreads();
        if (cls >= 0) {
            const int32_t pos = lbase[bin] + off;
            const int32_t a0 = arst[i], alen = arst[i + 1] - a0;
            for (int32_t t = 0; t < nt; ++t) {
                Task tk;
                tk.row = (int32_t)i;
                tk.a0 = a0 + (int32_t)((int64_t)alen * t / nt);
                tk.a1 = a0 + (int32_t)((int64_t)alen * (t + 1) / nt);
                tk.split = nt > 1;
                tasks[pos + t] = tk;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// task engine

// LDS prefix of B-row lengths for A entries [a0, a1) of one task (a wave or
// a tile of `G` lanes).  pre[x] = streamed entries before A entry a0+x;
// bst[x] = index in B.col of streamed entry 0 of A entry a0+x minus pre[x].
// Returns the number of streamed entries.
template <int G>
__device__ __forceinline__ int32_t build_prefix(const Prob &P, int32_t a0, int32_t a1, int32_t *pre, int32_t *bst,
                                                int gl) {
    int32_t total = 0;
    for (int32_t base = a0; base < a1; base += G) {
        int32_t p = base + gl;
        int32_t len = 0, bs = 0;
        if (p < a1) {
            int32_t k = P.acol[p];
            bs = P.brst[k];
            len = P.brst[k + 1] - bs;
        }
        int32_t incl = len;
#pragma unroll
        for (int o = 1; o < G; o <<= 1) {
            int32_t w = __shfl_up(incl, o, G);
            if (gl >= o) incl += w;
        }
        if (p < a1) {
            pre[p - a0] = total + incl - len;
            bst[p - a0] = bs - (total + incl - len);
        }
        total += __shfl(incl, G - 1, G);
    }
    if (gl == 0) pre[a1 - a0] = total;
    return total;
}

// Hit handling common to every accumulator.  q = mask position (0-based in
// the row), kk = A entry offset in the task, t = index of the streamed entry
// in B.col.
template <int MODE>
struct Hit;
// Modes with per-mask-position accumulators.
template <int MODE>
struct HasAcc {
    static constexpr bool v = MODE == MODE_VALUES || MODE == MODE_SUPPORT;
};

template <>
struct Hit<MODE_TC> {
    __device__ static void on(const Prob &, int32_t *, int32_t *, int32_t, int32_t, int32_t, int32_t, int32_t,
                              uint32_t &cnt) {
        ++cnt;
    }
};
template <>
struct Hit<MODE_VALUES> {
    __device__ static void on(const Prob &P, int32_t *acc, int32_t *, int32_t q, int32_t, int32_t a0, int32_t kk,
                              int32_t t, uint32_t &cnt) {
        ++cnt;
        uint32_t av = P.aval ? (uint32_t)P.aval[a0 + kk] : 1u;
        uint32_t bv = P.bval ? (uint32_t)P.bval[t] : 1u;
        atomicAdd((uint32_t *)&acc[q], av * bv);
    }
};
template <>
struct Hit<MODE_SUPPORT> {
    __device__ static void on(const Prob &P, int32_t *acc, int32_t *acnt, int32_t q, int32_t, int32_t, int32_t kk,
                              int32_t t, uint32_t &cnt) {
        ++cnt;
        atomicAdd((uint32_t *)&acc[q], 1u);          // edge (i, j)
        atomicAdd((uint32_t *)&acnt[kk], 1u);        // edge (i, k)
        atomicAdd((uint32_t *)&P.cval[t], 1u);       // edge (k, j)
    }
};

// Write back the per-task accumulators.
template <int MODE, int G>
__device__ __forceinline__ void flush(const Prob &P, const Task &tk, int32_t mb, int32_t mlen, const int32_t *acc,
                                      const int32_t *acnt, int gl) {
    if (MODE == MODE_VALUES) {
        for (int32_t q = gl; q < mlen; q += G) {
            if (tk.split) atomicAdd((uint32_t *)&P.cval[mb + q], (uint32_t)acc[q]);
            else P.cval[mb + q] = acc[q];
        }
    } else if (MODE == MODE_SUPPORT) {
        for (int32_t q = gl; q < mlen; q += G)
            if (acc[q]) atomicAdd((uint32_t *)&P.cval[mb + q], (uint32_t)acc[q]);
        for (int32_t x = gl; x < tk.a1 - tk.a0; x += G)
            if (acnt[x]) atomicAdd((uint32_t *)&P.cval[tk.a0 + x], (uint32_t)acnt[x]);
    }
}

template <int MODE>
__device__ __forceinline__ void tc_finish(const Prob &P, const Task &tk, uint32_t cnt_task, int gl, int G) {
    (void)gl; (void)G;
    if (MODE == MODE_TC && P.rowcount && cnt_task) atomicAdd(&P.rowcount[tk.row], (unsigned long long)cnt_task);
}

// A lane's share of a stream of `total` entries cut into rows by the
// prefix pre[0, na] (pre[0] = 0, pre[na] = total): a tile of G <= 64 lanes
// strides over the whole stream; a block of G > 64 lanes gives each wave a
// contiguous segment, so that a lane crosses the B rows of its wave's
// segment only (64 / row length prefix steps per entry, not G / row
// length).  first: the lane's first position; end: one past the segment;
// kk: the prefix row that holds the segment's start.
struct LaneRange {
    int32_t first, end, kk;
};
template <int G>
__device__ __forceinline__ LaneRange lane_range(const int32_t *pre, int32_t na, int32_t total, int gl) {
    LaneRange r{gl, total, 0};
    if constexpr (G > 64) {
        constexpr int NWAVE = G / 64;
        const int32_t per = (total + NWAVE - 1) / NWAVE;
        const int32_t s0 = min(total, (gl >> 6) * per);
        r.end = min(total, s0 + per);
        r.first = s0 + (gl & 63);
        int32_t lo = 0, hi = na - 1;  // the largest row with pre[row] <= s0
        while (lo < hi) {
            const int32_t mid = (lo + hi + 1) >> 1;
            if (pre[mid] <= s0) lo = mid; else hi = mid - 1;
        }
        r.kk = lo;
    }
    return r;
}

// Stream the concatenated B rows of a task with a tile of G lanes, U
// entries per lane per step (lane_range: blocks split it by waves): the
// cursor (A entry kk, its B-row base) lives in registers and is advanced
// through the LDS prefix only when an entry crosses into the next B row;
// the U loads of B columns are issued before any of them is probed.
// probe(b) returns the mask position of column b or -1.  Returns the number
// of hits.  CLAMP: positions past the lane range's end are clamped to its
// last entry, so that the walk and the loads need no lane branch and only
// the probe is skipped (faster where the streams are long; where many are
// short, the branches skip whole steps of loads).
template <int MODE, int G, int U, bool CLAMP = false, class Probe>
__device__ __forceinline__ uint32_t stream_entries(const Prob &P, const Task &tk, int32_t mb, const int32_t *pre,
                                                   const int32_t *bst, int32_t total, int gl, int32_t *acc,
                                                   int32_t *acnt, Probe probe) {
    uint32_t cnt = 0;
    if (total <= 0) return 0;
    constexpr int L = G > 64 ? 64 : G;
    const LaneRange r = lane_range<G>(pre, tk.a1 - tk.a0, total, gl);
    const int32_t end = r.end;
    int32_t kk = r.kk, nxt = pre[kk + 1], base = bst[kk];
    if constexpr (CLAMP) {
        for (int32_t g0 = r.first; g0 < end; g0 += L * U) {
            int32_t t[U], k[U], b[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int32_t g = min(g0 + L * u, end - 1);
                while (g >= nxt) {
                    ++kk;
                    nxt = pre[kk + 1];
                    base = bst[kk];
                }
                t[u] = base + g;
                k[u] = kk;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) b[u] = P.bcol[t[u]];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (g0 + L * u < end) {
                    const int32_t q = probe(b[u]);
                    if (q >= 0) Hit<MODE>::on(P, acc, acnt, q, mb, tk.a0, k[u], t[u], cnt);
                }
            }
        }
        return cnt;
    }
    for (int32_t g0 = r.first; g0 < end; g0 += L * U) {
        int32_t t[U], k[U], b[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int32_t g = g0 + L * u;
            t[u] = -1;
            k[u] = 0;
            if (g < end) {
                while (g >= nxt) {
                    ++kk;
                    nxt = pre[kk + 1];
                    base = bst[kk];
                }
                t[u] = base + g;
                k[u] = kk;
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) b[u] = t[u] >= 0 ? P.bcol[t[u]] : -1;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (t[u] >= 0) {
                const int32_t q = probe(b[u]);
                if (q >= 0) Hit<MODE>::on(P, acc, acnt, q, mb, tk.a0, k[u], t[u], cnt);
            }
        }
    }
    return cnt;
}

// Triangle counting over the hub rows of a task: the concatenated bitmap
// words of its hub B rows (hpre / hbst from WinPrefix), U words per
// lane in flight; word w of row k holds columns [32w, 32w + 32), all below
// k, and is ANDed with word w of the mask row's LDS bitmap.  Consecutive
// lanes read consecutive words of a row.  Returns the number of hits.
template <int G, int U>
__device__ __forceinline__ uint32_t stream_hub(const Prob &P, const int32_t *hpre, const int32_t *hbst,
                                               int32_t na, int32_t total, int gl, const uint32_t *bm) {
    uint32_t cnt = 0;
    if (total <= 0) return 0;
    constexpr int L = G > 64 ? 64 : G;
    const LaneRange r = lane_range<G>(hpre, na, total, gl);
    const int32_t end = r.end;
    int32_t kk = r.kk, cur = hpre[kk], nxt = hpre[kk + 1], base = hbst[kk];
    for (int32_t g0 = r.first; g0 < end; g0 += L * U) {
        int32_t t[U], w[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int32_t g = g0 + L * u;
            t[u] = -1;
            w[u] = 0;
            if (g < end) {
                while (g >= nxt) {
                    ++kk;
                    cur = nxt;
                    nxt = hpre[kk + 1];
                    base = hbst[kk];
                }
                t[u] = base + g;
                w[u] = g - cur;
            }
        }
        uint32_t v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = t[u] >= 0 ? P.hub[t[u]] : 0u;
#pragma unroll
        for (int u = 0; u < U; ++u) cnt += __popc(v[u] & bm[w[u]]);
    }
    return cnt;
}

// Triangle counting by bit tests (hub_choice): test j of A entry x is the
// mask row's column c = mcol[mb + j] (below the entry's column k) against
// bit c of hub row k at word offset pofs[x]; U tests per lane in flight.
// Returns the number of hits.
template <int G, int U>
__device__ __forceinline__ uint32_t stream_probe(const Prob &P, const int32_t *ppre, const int32_t *pofs,
                                                 int32_t na, int32_t total, int gl, int32_t mb) {
    uint32_t cnt = 0;
    if (total <= 0) return 0;
    constexpr int L = G > 64 ? 64 : G;
    const LaneRange r = lane_range<G>(ppre, na, total, gl);
    const int32_t end = r.end;
    int32_t kk = r.kk, cur = ppre[kk], nxt = ppre[kk + 1], base = pofs[kk];
    for (int32_t g0 = r.first; g0 < end; g0 += L * U) {
        int32_t off[U], c[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int32_t g = g0 + L * u;
            off[u] = -1;
            c[u] = 0;
            if (g < end) {
                while (g >= nxt) {
                    ++kk;
                    cur = nxt;
                    nxt = ppre[kk + 1];
                    base = pofs[kk];
                }
                off[u] = base;
                c[u] = P.mcol[mb + (g - cur)];
            }
        }
        uint32_t v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = off[u] >= 0 ? P.hub[off[u] + (c[u] >> 5)] : 0u;
#pragma unroll
        for (int u = 0; u < U; ++u) cnt += (v[u] >> (c[u] & 31)) & 1u;
    }
    return cnt;
}

template <int MODE>
__device__ __forceinline__ void finish_kernel(const Prob &P, uint32_t cnt, int cls, int lane) {
    unsigned long long s = dev::wave_sum((unsigned long long)cnt);
    if (lane == 0 && s) {
        if (MODE == MODE_TC) atomicAdd(P.count, s);
        atomicAdd(&P.hits[cls], s);
    }
}

// ---- MSA-LDS --------------------------------------------------------------
// One wave per task; LDS: key[HC] (mask column or -1), pos[HC] (mask
// position), acc[MAXM] (per-position counters), pre/bst[kMaxTaskA+1].  Two
// instantiations: a small table for short mask rows (high occupancy) and a
// large one.  A task's table has the smallest power-of-two capacity >= 4 x
// its mask length (at most HC), so probes rarely walk past one slot.
template <int MODE, int HC, int MAXM>
__global__ __launch_bounds__(64) void k_msa_lds(Prob P, int cls) {
    __shared__ int32_t hkey[HC];
    __shared__ int32_t hpos[MODE == MODE_TC ? 1 : HC];
    __shared__ int32_t acc[HasAcc<MODE>::v ? MAXM : 1];
    __shared__ int32_t acnt[MODE == MODE_SUPPORT ? kMaxTaskA : 1];
    __shared__ int32_t pre[kMaxTaskA + 1];
    __shared__ int32_t bst[kMaxTaskA];
    const int lane = threadIdx.x;
    const int32_t lo = P.bin_off[cls * kFlopBins], hi = P.bin_off[(cls + 1) * kFlopBins];
    uint32_t cnt = 0;
    for (int32_t ti = lo + blockIdx.x; ti < hi; ti += gridDim.x) {
        const Task tk = P.tasks[ti];
        const long long tstart = P.task_cycles ? clock64() : 0;
        const int32_t mb = P.mrst[tk.row], mlen = P.mrst[tk.row + 1] - mb;
        int lg = 6;
        while ((1 << lg) < 4 * mlen && (1 << lg) < HC) ++lg;
        const uint32_t cmask = (1u << lg) - 1;
        for (uint32_t s = lane; s <= cmask; s += 64) hkey[s] = -1;
        if (HasAcc<MODE>::v)
            for (int32_t q = lane; q < mlen; q += 64) acc[q] = 0;
        if (MODE == MODE_SUPPORT)
            for (int32_t x = lane; x < tk.a1 - tk.a0; x += 64) acnt[x] = 0;
        __syncthreads();
        for (int32_t q = lane; q < mlen; q += 64) {
            int32_t c = P.mcol[mb + q];
            uint32_t s = dev::hash32((uint32_t)c) >> (32 - lg);
            while (atomicCAS(&hkey[s], -1, c) != -1) s = (s + 1) & cmask;
            if (MODE != MODE_TC) hpos[s] = q;
        }
        const int32_t total = build_prefix<64>(P, tk.a0, tk.a1, pre, bst, lane);
        __syncthreads();
        const uint32_t c = stream_entries<MODE, 64, 4>(P, tk, mb, pre, bst, total, lane, acc, acnt,
                                                       [&](int32_t b) -> int32_t {
            uint32_t s = dev::hash32((uint32_t)b) >> (32 - lg);
            for (;;) {
                const int32_t key = hkey[s];
                if (key == b) return MODE == MODE_TC ? 0 : hpos[s];
                if (key < 0) return -1;
                s = (s + 1) & cmask;
            }
        });
        __syncthreads();
        flush<MODE, 64>(P, tk, mb, mlen, acc, acnt, lane);
        if (MODE == MODE_TC && P.rowcount) {
            uint32_t r = dev::wave_sum(c);
            if (lane == 0 && r) atomicAdd(&P.rowcount[tk.row], (unsigned long long)r);
        }
        cnt += c;
        if (P.task_cycles && threadIdx.x == 0) P.task_cycles[ti] = clock64() - tstart;
        __syncthreads();
    }
    finish_kernel<MODE>(P, cnt, cls, lane);
}

// ---- MSA-LDS, dense window (triangle counting) ----------------------------
// With vertices numbered by descending degree, the columns streamed by
// C<L> = L*L concentrate on the hubs: on R-MAT scale 22, 98.8% of them are
// below 32768.  This kernel keeps the mask row's columns below kWindow as a
// dense bitmap in LDS (one bit per column, the MSA dense accumulator over a
// window) and the few columns at or above kWindow in a small LDS hash, so
// almost every probe is one LDS word read and a bit test.  Only the bitmap
// words a row sets are cleared afterwards.

// The B-row prefix of one window task (up to kMaxTaskA A entries), built in
// two steps so that every global load of a task's set-up is issued before
// the first barrier: load() reads the A entries' columns and B row pointers
// (and picks each hub row's way, hub_choice); a one-wave block then writes
// the whole prefix (scan_wave), a larger block (one A entry per thread)
// scans per wave before the barrier (scan_block1) and adds the wave totals
// after it (scan_block2).  pre/bst as build_prefix; hpre[x] = bitmap words
// before A entry a0+x, hbst[x] = its row's word offset in P.hub minus
// hpre[x]; ppre[x] = bit tests before it, pofs[x] = its row's word offset.
template <int NT, bool HUB>
struct WinPrefix {
    static constexpr int R = (kMaxTaskA + NT - 1) / NT;
    static constexpr int NW = NT / 64;
    static_assert(NT == 64 || R == 1, "one-wave blocks or one A entry per thread");
    int32_t len[R], bs[R], hlen[R], hbs[R], plen[R], pbs[R];
    int32_t incl = 0, hincl = 0, pincl = 0;  // scan_block1: the thread's wave-inclusive sums
    int32_t total = 0, htotal = 0, ptotal = 0;

    __device__ __forceinline__ void load(const Prob &P, int32_t a0, int32_t a1, int32_t mb) {
        const int tid = threadIdx.x;
        int32_t k[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int32_t p = a0 + r * NT + tid;
            k[r] = p < a1 ? P.acol[p] : -1;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            bs[r] = k[r] >= 0 ? P.brst[k[r]] : 0;
            len[r] = k[r] >= 0 ? P.brst[k[r] + 1] : 0;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            len[r] -= bs[r];
            hlen[r] = hbs[r] = plen[r] = pbs[r] = 0;
            if (HUB && k[r] >= 0) hub_choice(P, k[r], a0 + r * NT + tid - mb, len[r], hlen[r], hbs[r], plen[r], pbs[r]);
        }
    }

    __device__ __forceinline__ void scan_wave(int32_t a0, int32_t a1, int32_t *pre, int32_t *bst, int32_t *hpre,
                                              int32_t *hbst, int32_t *ppre, int32_t *pofs) {
        static_assert(NT == 64, "one-wave blocks");
        const int lane = threadIdx.x;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (a0 + r * 64 >= a1) break;
            const int32_t p = a0 + r * 64 + lane;
            const int32_t i1 = dev::wave_incl_scan(len[r]);
            const int32_t i2 = HUB ? dev::wave_incl_scan(hlen[r]) : 0;
            const int32_t i3 = HUB ? dev::wave_incl_scan(plen[r]) : 0;
            if (p < a1) {
                const int32_t ex = total + i1 - len[r];
                pre[p - a0] = ex;
                bst[p - a0] = bs[r] - ex;
                if (HUB) {
                    const int32_t hex = htotal + i2 - hlen[r];
                    hpre[p - a0] = hex;
                    hbst[p - a0] = hbs[r] - hex;
                    ppre[p - a0] = ptotal + i3 - plen[r];
                    pofs[p - a0] = pbs[r];
                }
            }
            total += __shfl(i1, 63);
            if (HUB) {
                htotal += __shfl(i2, 63);
                ptotal += __shfl(i3, 63);
            }
        }
        if (lane == 0) {
            pre[a1 - a0] = total;
            if (HUB) {
                hpre[a1 - a0] = htotal;
                ppre[a1 - a0] = ptotal;
            }
        }
    }

    // wsum: 3 * NW ints
    __device__ __forceinline__ void scan_block1(int32_t *wsum) {
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        incl = dev::wave_incl_scan(len[0]);
        if (HUB) {
            hincl = dev::wave_incl_scan(hlen[0]);
            pincl = dev::wave_incl_scan(plen[0]);
        }
        if (lane == 63) {
            wsum[w] = incl;
            if (HUB) {
                wsum[NW + w] = hincl;
                wsum[2 * NW + w] = pincl;
            }
        }
    }

    __device__ __forceinline__ void scan_block2(int32_t a0, int32_t a1, int32_t *pre, int32_t *bst, int32_t *hpre,
                                                int32_t *hbst, int32_t *ppre, int32_t *pofs, const int32_t *wsum) {
        const int tid = threadIdx.x, w = tid >> 6;
        int32_t before = 0, hbefore = 0, pbefore = 0;
#pragma unroll
        for (int x = 0; x < NW; ++x) {
            const int32_t v = wsum[x];
            before += x < w ? v : 0;
            total += v;
            if (HUB) {
                const int32_t hv = wsum[NW + x], pv = wsum[2 * NW + x];
                hbefore += x < w ? hv : 0;
                htotal += hv;
                pbefore += x < w ? pv : 0;
                ptotal += pv;
            }
        }
        const int32_t p = a0 + tid;
        if (p < a1) {
            const int32_t ex = before + incl - len[0];
            pre[p - a0] = ex;
            bst[p - a0] = bs[0] - ex;
            if (HUB) {
                const int32_t hex = hbefore + hincl - hlen[0];
                hpre[p - a0] = hex;
                hbst[p - a0] = hbs[0] - hex;
                ppre[p - a0] = pbefore + pincl - plen[0];
                pofs[p - a0] = pbs[0];
            }
        }
        if (tid == 0) {
            pre[a1 - a0] = total;
            if (HUB) {
                hpre[a1 - a0] = htotal;
                ppre[a1 - a0] = ptotal;
            }
        }
    }
};

// TC and SUPPORT modes.  SUPPORT also needs the mask position of a hit:
// the sorted mask row is copied to LDS and a hit below kWindow finds its
// position by binary search there (hits are a few percent of the probes);
// the hash for the columns above the window stores positions directly.
// A task's set-up is one round of global loads: the mask row (held in
// registers until the task ends, when its bitmap words are cleared from
// them), the A columns and B row pointers (WinPrefix), all issued before
// the first barrier.  Blocks take tasks by ticket (P.ticket), K consecutive
// tasks per ticket (ticket_chunk), the next ticket while the chunk's last
// task streams, so a class whose blocks outnumber the slots free beside the
// other classes spreads onto the CUs they release.
// The hash starts empty for every
// task (it is cleared after a task that used it), with a capacity fixed by
// the mask row's length, so columns go into it as they are read.
template <int MODE, int NW, int HC, int MAXM, bool HUB>
__global__ __launch_bounds__(64 * NW) void k_msa_window(Prob P, int cls) {
    static_assert(MODE == MODE_TC || MODE == MODE_SUPPORT, "window kernel modes");
    constexpr int NT = 64 * NW;
    constexpr int CPT = (MAXM + NT - 1) / NT;  // mask columns per thread
    constexpr bool SUP = MODE == MODE_SUPPORT;
    __shared__ uint32_t bm[kWindow / 32];
    __shared__ int32_t hkey[HC];
    __shared__ int32_t hpos[SUP ? HC : 1];
    __shared__ int32_t mk[SUP ? MAXM : 1];
    __shared__ int32_t acc[SUP ? MAXM : 1];
    __shared__ int32_t acnt[SUP ? kMaxTaskA : 1];
    __shared__ int32_t pre[kMaxTaskA + 1];
    __shared__ int32_t bst[kMaxTaskA];
    // TC with hub bitmaps: prefix of the hub rows' bitmap words
    static_assert(!HUB || !SUP, "hub bitmaps count triangles only");
    __shared__ int32_t hpre[HUB ? kMaxTaskA + 1 : 1];
    __shared__ int32_t hbst[HUB ? kMaxTaskA : 1];
    __shared__ int32_t ppre[HUB ? kMaxTaskA + 1 : 1];
    __shared__ int32_t pofs[HUB ? kMaxTaskA : 1];
    __shared__ int32_t wsum[3 * NW];
    __shared__ int32_t nhigh_s;  // the task's mask columns >= kWindow
    __shared__ int32_t next_s;   // the block's next task
    const int tid = threadIdx.x;
    const int32_t lo = P.bin_off[cls * kFlopBins], hi = P.bin_off[(cls + 1) * kFlopBins];
    const int32_t K = ticket_chunk(hi - lo, (int32_t)gridDim.x, 1);
    for (int w = tid; w < kWindow / 32; w += NT) bm[w] = 0;
    for (int s = tid; s < HC; s += NT) hkey[s] = -1;
    if (tid == 0) {
        nhigh_s = 0;
        next_s = lo + atomicAdd(&P.ticket[cls], 1) * K;
    }
    __syncthreads();
    uint32_t cnt = 0;
    for (int32_t ti = __builtin_amdgcn_readfirstlane(next_s); ti < hi;
         ti = __builtin_amdgcn_readfirstlane(next_s)) {
        const long long tstart = P.task_cycles ? clock64() : 0;
        const Task tk = P.tasks[ti];
        const int32_t mb = P.mrst[tk.row], mlen = P.mrst[tk.row + 1] - mb;
        int lg = 6;
        while ((1 << lg) < 4 * mlen && (1 << lg) < HC) ++lg;
        const uint32_t cmask = (1u << lg) - 1;
        int32_t mc[CPT];
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const int32_t q = tid + j * NT;
            mc[j] = q < mlen ? P.mcol[mb + q] : -1;
        }
        WinPrefix<NT, HUB> pf;
        pf.load(P, tk.a0, tk.a1, mb);
        if (SUP) {
            for (int32_t q = tid; q < mlen; q += NT) acc[q] = 0;
            for (int32_t x = tid; x < tk.a1 - tk.a0; x += NT) acnt[x] = 0;
        }
        // mask columns are sorted: the ones below kWindow are positions
        // [0, nlow) and go to the bitmap, the rest to the hash
        int32_t nh = 0;
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const int32_t q = tid + j * NT, c = mc[j];
            if (c < 0) continue;
            if (c < kWindow) {
                atomicOr(&bm[c >> 5], 1u << (c & 31));
                if (SUP) mk[q] = c;
            } else {
                uint32_t s = dev::hash32((uint32_t)c) >> (32 - lg);
                while (atomicCAS(&hkey[s], -1, c) != -1) s = (s + 1) & cmask;
                if (SUP) hpos[s] = q;
                ++nh;
            }
        }
        if (nh) atomicAdd(&nhigh_s, nh);
        if constexpr (NW == 1) pf.scan_wave(tk.a0, tk.a1, pre, bst, hpre, hbst, ppre, pofs);
        else pf.scan_block1(wsum);
        __syncthreads();
        const int32_t nhigh = nhigh_s, nlow = mlen - nhigh;
        // the next chunk's ticket, taken while the chunk's last task streams
        const bool chunk_end = (ti - lo) % K == K - 1;
        int32_t tick = 0;
        if (tid == 0 && chunk_end) tick = atomicAdd(&P.ticket[cls], 1);
        if constexpr (NW > 1) {
            pf.scan_block2(tk.a0, tk.a1, pre, bst, hpre, hbst, ppre, pofs, wsum);
            __syncthreads();
        }
        uint32_t c = 0;
        if constexpr (HUB) {
            if (pf.htotal) c = stream_hub<NT, kWindowU>(P, hpre, hbst, tk.a1 - tk.a0, pf.htotal, tid, bm);
            if (pf.ptotal) c += stream_probe<NT, kWindowU>(P, ppre, pofs, tk.a1 - tk.a0, pf.ptotal, tid, mb);
        }
        // the window kernels without hub bitmaps stream long B rows: clamped
        c += stream_entries<MODE, NT, kWindowU, !HUB>(P, tk, mb, pre, bst, pf.total, tid, acc, acnt,
                                                       [&](int32_t b) -> int32_t {
            if (b < kWindow) {
                if (!((bm[b >> 5] >> (b & 31)) & 1u)) return -1;
                if (!SUP) return 0;
                int32_t x = 0, y = nlow;  // position of b in mk[0, nlow)
                while (x < y) {
                    int32_t mid = x + ((y - x) >> 1);
                    if (mk[mid] < b) x = mid + 1; else y = mid;
                }
                return x;
            }
            if (!nhigh) return -1;
            uint32_t s = dev::hash32((uint32_t)b) >> (32 - lg);
            for (;;) {
                const int32_t key = hkey[s];
                if (key == b) return SUP ? hpos[s] : 0;
                if (key < 0) return -1;
                s = (s + 1) & cmask;
            }
        });
        if (tid == 0) next_s = chunk_end ? lo + tick * K : ti + 1;
        __syncthreads();
#pragma unroll
        for (int j = 0; j < CPT; ++j)
            if (mc[j] >= 0 && mc[j] < kWindow) bm[mc[j] >> 5] = 0;
        if (nhigh)
            for (uint32_t s = tid; s <= cmask; s += NT) hkey[s] = -1;
        if (tid == 0) nhigh_s = 0;
        if (SUP) flush<MODE, NT>(P, tk, mb, mlen, acc, acnt, tid);
        if (MODE == MODE_TC && P.rowcount) {
            uint32_t r = dev::wave_sum(c);
            if ((tid & 63) == 0 && r) atomicAdd(&P.rowcount[tk.row], (unsigned long long)r);
        }
        cnt += c;
        if (P.task_cycles && threadIdx.x == 0) P.task_cycles[ti] = clock64() - tstart;
        __syncthreads();
    }
    finish_kernel<MODE>(P, cnt, cls, tid & 63);
}

// ---- MCA-wave -------------------------------------------------------------
// Triangle counting with hub bitmaps: an A entry k below hub_n whose B row
// has more entries than the mask row has columns below k is counted by
// testing those columns' bits in row k's bitmap instead of streaming the
// row.  pre/bst as build_prefix for the streamed entries; hpre[x] = bit
// tests before A entry a0+x, hbst[x] = word offset of its bitmap row.
// m[0, mlen) is the tile's sorted mask row.  Returns (streamed, tests).
template <int G>
__device__ __forceinline__ int2 build_prefix_probe(const Prob &P, int32_t a0, int32_t a1, const int32_t *m,
                                                   int32_t mlen, int32_t *pre, int32_t *bst, int32_t *hpre,
                                                   int32_t *hbst, int gl) {
    int32_t total = 0, htotal = 0;
    for (int32_t base = a0; base < a1; base += G) {
        const int32_t p = base + gl;
        int32_t len = 0, bs = 0, hlen = 0, hbs = 0;
        if (p < a1) {
            const int32_t k = P.acol[p];
            bs = P.brst[k];
            len = P.brst[k + 1] - bs;
            if (k < P.hub_n && len > 0) {
                int32_t l = 0, h = mlen;  // mask columns below k
                while (l < h) {
                    const int32_t mid = l + ((h - l) >> 1);
                    if (m[mid] < k) l = mid + 1; else h = mid;
                }
                if (l < len) {
                    hlen = l;
                    hbs = k * P.hub_w;
                    len = 0;
                }
            }
        }
        int32_t incl = len, hincl = hlen;
#pragma unroll
        for (int o = 1; o < G; o <<= 1) {
            const int32_t w = __shfl_up(incl, o, G), hw = __shfl_up(hincl, o, G);
            if (gl >= o) {
                incl += w;
                hincl += hw;
            }
        }
        if (p < a1) {
            pre[p - a0] = total + incl - len;
            bst[p - a0] = bs - (total + incl - len);
            hpre[p - a0] = htotal + hincl - hlen;
            hbst[p - a0] = hbs;
        }
        total += __shfl(incl, G - 1, G);
        htotal += __shfl(hincl, G - 1, G);
    }
    if (gl == 0) {
        pre[a1 - a0] = total;
        hpre[a1 - a0] = htotal;
    }
    return make_int2(total, htotal);
}

// Filter slot of column c in an MCA-wave tile's mask filter (multiplicative
// hash, kMcaFilterBits slots).
__device__ __forceinline__ uint32_t mca_hash(int32_t c) {
    return ((uint32_t)c * 2654435761u) >> (32 - kMcaFilterLg);
}

// Tiles of G lanes, 64/G tasks per wave.  The sorted mask row sits in LDS;
// every lane binary-searches its streamed B entry in it (after a range
// check against the row's first and last column).  In TC mode a bit filter
// of the mask row comes first: only entries whose bit is set are searched
// (one per lane per pass), and each lane counts its own hits.
template <int MODE, int G, bool HUB>
__global__ __launch_bounds__(64) void k_mca_wave(Prob P, int cls) {
    constexpr int T = 64 / G;
    __shared__ int32_t mk[T][kMcaMaxMask];
    __shared__ int32_t acc[T][HasAcc<MODE>::v ? kMcaMaxMask : 1];
    __shared__ int32_t acnt[T][MODE == MODE_SUPPORT ? kMcaMaxTaskA : 1];
    __shared__ int32_t pre[T][kMcaMaxTaskA + 1];
    __shared__ int32_t bst[T][kMcaMaxTaskA];
    static_assert(!HUB || MODE == MODE_TC, "hub bitmaps count triangles only");
    constexpr bool TCH = HUB;
    __shared__ int32_t hpre[TCH ? T : 1][kMcaMaxTaskA + 1];
    __shared__ int32_t hbst[TCH ? T : 1][kMcaMaxTaskA];
    // TC and supports: a bit filter of each tile's mask row (bit mca_hash(c)
    // per column)
    constexpr bool FLT = MODE == MODE_TC || MODE == MODE_SUPPORT;
    __shared__ uint32_t flt[FLT ? T : 1][FLT ? kMcaFilterWords : 1];
    const int lane = threadIdx.x, tile = lane / G, gl = lane % G;
    if constexpr (FLT) {
        for (int w = gl; w < kMcaFilterWords; w += G) flt[tile][w] = 0;
        __syncthreads();
    }
    const int32_t lo = P.bin_off[cls * kFlopBins], hi = P.bin_off[(cls + 1) * kFlopBins];
    uint32_t cnt = 0;
    // The wave takes RB >= kMcaRounds rounds of T tasks per ticket
    // (P.ticket, ticket_chunk): tasks here are small and many, and one
    // ticket per round would serialise on the shared counter.  The next
    // ticket is taken during a chunk's last round.  All tiles run the same
    // rounds, so the wave stays converged at the barriers.
    const int32_t RB = ticket_chunk((hi - lo) / T, (int32_t)gridDim.x, kMcaRounds);
    int32_t first = 0, tick = 0, r = 0;
    if (lane == 0) first = atomicAdd(&P.ticket[cls], T * RB);
    first = lo + __builtin_amdgcn_readfirstlane(first);
    while (first + r * T < hi) {
        const int32_t ti = first + r * T + tile;
        const bool live = ti < hi;
        Task tk{0, 0, 0, 0};
        int32_t mb = 0, mlen = 0;
        const long long tstart = P.task_cycles ? clock64() : 0;
        if (live) {
            tk = P.tasks[ti];
            mb = P.mrst[tk.row];
            mlen = P.mrst[tk.row + 1] - mb;
        }
        for (int32_t q = gl; q < mlen; q += G) {
            const int32_t c = P.mcol[mb + q];
            mk[tile][q] = c;
            if (HasAcc<MODE>::v) acc[tile][q] = 0;
            if constexpr (FLT) {
                const uint32_t h = mca_hash(c);
                atomicOr(&flt[tile][h >> 5], 1u << (h & 31));
            }
        }
        if (MODE == MODE_SUPPORT)
            for (int32_t x = gl; x < tk.a1 - tk.a0; x += G) acnt[tile][x] = 0;
        int32_t total = 0, htotal = 0;
        if constexpr (HUB) {
            __syncthreads();  // the mask rows, for the bit-test counts
            if (live) {
                const int2 tt = build_prefix_probe<G>(P, tk.a0, tk.a1, mk[tile], mlen, pre[tile], bst[tile],
                                                      hpre[TCH ? tile : 0], hbst[TCH ? tile : 0], gl);
                total = tt.x;
                htotal = tt.y;
            }
        } else if (live) {
            total = build_prefix<G>(P, tk.a0, tk.a1, pre[tile], bst[tile], gl);
        }
        __syncthreads();
        if (r == RB - 1 && lane == 0) tick = atomicAdd(&P.ticket[cls], T * RB);
        uint32_t cnt_task = 0;
        if (live && mlen > 0 && total > 0) {
            const int32_t *m = mk[tile];
            const int32_t mlo = m[0], mhi = m[mlen - 1];
            int32_t kk = 0, nxt = pre[tile][1], base = bst[tile][0];
            // kMcaU streamed entries per lane in flight
            for (int32_t g0 = gl; g0 < total; g0 += G * kMcaU) {
                int32_t t[kMcaU], kx[kMcaU], b[kMcaU];
#pragma unroll
                for (int u = 0; u < kMcaU; ++u) {
                    const int32_t g = g0 + G * u;
                    t[u] = -1;
                    kx[u] = 0;
                    if (g < total) {
                        while (g >= nxt) {
                            ++kk;
                            nxt = pre[tile][kk + 1];
                            base = bst[tile][kk];
                        }
                        t[u] = base + g;
                        kx[u] = kk;
                    }
                }
#pragma unroll
                for (int u = 0; u < kMcaU; ++u) b[u] = t[u] >= 0 ? P.bcol[t[u]] : -1;
                if constexpr (FLT) {
                    // Entries whose filter bit is clear are not in the mask
                    // row (almost all of them: hits are rare); a lane
                    // binary-searches one remaining candidate per pass, so
                    // the wave runs about one search per batch of kMcaU
                    // entries instead of one per entry
                    uint32_t cand = 0;
#pragma unroll
                    for (int u = 0; u < kMcaU; ++u) {
                        const uint32_t h = mca_hash(b[u]);
                        const bool in = t[u] >= 0 && b[u] >= mlo && b[u] <= mhi &&
                                        ((flt[tile][h >> 5] >> (h & 31)) & 1u);
                        cand |= in ? 1u << u : 0u;
                    }
                    while (__any(cand != 0)) {
                        int32_t bb = -1, tt = 0, kq = 0;
                        if (cand) {
                            const int uu = __builtin_ctz(cand);
                            cand &= cand - 1;
#pragma unroll
                            for (int u = 0; u < kMcaU; ++u) {
                                bb = u == uu ? b[u] : bb;
                                tt = u == uu ? t[u] : tt;
                                kq = u == uu ? kx[u] : kq;
                            }
                        }
                        if (bb >= 0) {
                            int32_t l = 0, h = mlen;  // lower_bound in m[0, mlen)
                            while (l < h) {
                                const int32_t mid = l + ((h - l) >> 1);
                                if (m[mid] < bb) l = mid + 1; else h = mid;
                            }
                            if (m[l] == bb) {
                                if (MODE == MODE_TC) ++cnt_task;
                                else Hit<MODE>::on(P, acc[tile], acnt[tile], l, mb, tk.a0, kq, tt, cnt_task);
                            }
                        }
                    }
                    continue;
                }
#pragma unroll
                for (int u = 0; u < kMcaU; ++u) {
                    int32_t q = -1;
                    if (t[u] >= 0 && b[u] >= mlo && b[u] <= mhi) {
                        int32_t l = 0, h = mlen;  // lower_bound in m[0, mlen)
                        while (l < h) {
                            int32_t mid = l + ((h - l) >> 1);
                            if (m[mid] < b[u]) l = mid + 1; else h = mid;
                        }
                        if (m[l] == b[u]) q = l;
                    }
                    if (MODE == MODE_TC) {
                        const uint64_t found = __ballot(q >= 0);
                        if (gl == 0)
                            cnt_task += __popcll((found >> (tile * G)) & ((G == 64) ? ~0ull : ((1ull << G) - 1)));
                    } else if (q >= 0) {
                        Hit<MODE>::on(P, acc[tile], acnt[tile], q, mb, tk.a0, kx[u], t[u], cnt_task);
                    }
                }
            }
        }
        if (TCH && live && htotal > 0) {
            // bit tests: test g of A entry kk is mask column m[g - first test of kk]
            const int32_t *m = mk[tile];
            const int32_t *hp = hpre[TCH ? tile : 0], *hb = hbst[TCH ? tile : 0];
            int32_t kk = 0, cur = 0, nxt = hp[1], base = hb[0];
            for (int32_t g0 = gl; g0 < htotal; g0 += G * kMcaU) {
                int32_t a[kMcaU], bit[kMcaU];
#pragma unroll
                for (int u = 0; u < kMcaU; ++u) {
                    const int32_t g = g0 + G * u;
                    a[u] = -1;
                    bit[u] = 0;
                    if (g < htotal) {
                        while (g >= nxt) {
                            ++kk;
                            cur = nxt;
                            nxt = hp[kk + 1];
                            base = hb[kk];
                        }
                        const int32_t j = m[g - cur];
                        a[u] = base + (j >> 5);
                        bit[u] = j & 31;
                    }
                }
                uint32_t w[kMcaU];
#pragma unroll
                for (int u = 0; u < kMcaU; ++u) w[u] = a[u] >= 0 ? P.hub[a[u]] : 0u;
#pragma unroll
                for (int u = 0; u < kMcaU; ++u) cnt_task += (w[u] >> bit[u]) & 1u;
            }
        }
        __syncthreads();
        if constexpr (FLT)
            for (int32_t q = gl; q < mlen; q += G) flt[tile][mca_hash(mk[tile][q]) >> 5] = 0;
        if (live) {
            flush<MODE, G>(P, tk, mb, mlen, acc[tile], acnt[tile], gl);
            if (MODE == MODE_TC && P.rowcount) {
                uint32_t s = cnt_task;
#pragma unroll
                for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
                if (gl == 0 && s) atomicAdd(&P.rowcount[tk.row], (unsigned long long)s);
            }
            cnt += cnt_task;
            if (P.task_cycles && gl == 0) P.task_cycles[ti] = clock64() - tstart;
        }
        __syncthreads();
        if (++r == RB) {
            r = 0;
            first = lo + __builtin_amdgcn_readfirstlane(tick);
        }
    }
    finish_kernel<MODE>(P, cnt, cls, lane);
}

// ---- hash-global ----------------------------------------------------------
// One wave per task, its table (key, position) in HBM scratch of `cap`
// slots per wave; mask-position accumulators in HBM too.
template <int MODE>
__global__ __launch_bounds__(64) void k_hash_global(Prob P, int cls, int32_t *__restrict__ gkey,
                                                    int32_t *__restrict__ gpos, int32_t *__restrict__ gacc,
                                                    int lgcap) {
    __shared__ int32_t pre[kMaxTaskA + 1];
    __shared__ int32_t bst[kMaxTaskA];
    __shared__ int32_t acnt[MODE == MODE_SUPPORT ? kMaxTaskA : 1];
    const int lane = threadIdx.x;
    const uint32_t cap = 1u << lgcap;
    int32_t *key = gkey + (size_t)blockIdx.x * cap;
    int32_t *pos = gpos + (size_t)blockIdx.x * cap;
    int32_t *acc = gacc + (size_t)blockIdx.x * (cap >> 1);
    const int32_t lo = P.bin_off[cls * kFlopBins], hi = P.bin_off[(cls + 1) * kFlopBins];
    uint32_t cnt = 0;
    for (int32_t ti = lo + blockIdx.x; ti < hi; ti += gridDim.x) {
        const Task tk = P.tasks[ti];
        const long long tstart = P.task_cycles ? clock64() : 0;
        const int32_t mb = P.mrst[tk.row], mlen = P.mrst[tk.row + 1] - mb;
        int lg = 6;
        while ((1 << lg) < 2 * mlen && lg < lgcap) ++lg;
        const uint32_t cmask = (1u << lg) - 1;
        for (uint32_t s = lane; s <= cmask; s += 64) key[s] = -1;
        if (HasAcc<MODE>::v)
            for (int32_t q = lane; q < mlen; q += 64) acc[q] = 0;
        if (MODE == MODE_SUPPORT)
            for (int32_t x = lane; x < tk.a1 - tk.a0; x += 64) acnt[x] = 0;
        __threadfence_block();
        __syncthreads();
        for (int32_t q = lane; q < mlen; q += 64) {
            int32_t col = P.mcol[mb + q];
            uint32_t s = dev::hash32((uint32_t)col) >> (32 - lg);
            while (atomicCAS(&key[s], -1, col) != -1) s = (s + 1) & cmask;
            if (MODE != MODE_TC) pos[s] = q;
        }
        const int32_t total = build_prefix<64>(P, tk.a0, tk.a1, pre, bst, lane);
        __threadfence_block();
        __syncthreads();
        const uint32_t c = stream_entries<MODE, 64, 4>(P, tk, mb, pre, bst, total, lane, acc, acnt,
                                                       [&](int32_t b) -> int32_t {
            uint32_t s = dev::hash32((uint32_t)b) >> (32 - lg);
            for (;;) {
                const int32_t k2 = key[s];
                if (k2 == b) return MODE == MODE_TC ? 0 : pos[s];
                if (k2 < 0) return -1;
                s = (s + 1) & cmask;
            }
        });
        __threadfence_block();
        __syncthreads();
        flush<MODE, 64>(P, tk, mb, mlen, acc, acnt, lane);
        if (MODE == MODE_TC && P.rowcount) {
            uint32_t r = dev::wave_sum(c);
            if (lane == 0 && r) atomicAdd(&P.rowcount[tk.row], (unsigned long long)r);
        }
        cnt += c;
        if (P.task_cycles && threadIdx.x == 0) P.task_cycles[ti] = clock64() - tstart;
        __threadfence_block();
        __syncthreads();
    }
    finish_kernel<MODE>(P, cnt, cls, lane);
}

// ---- dense-global ---------------------------------------------------------
// One wave per task; a map of B.n entries per wave in HBM scratch holds
// mask position + 1 at each mask column (0 elsewhere).  Only the mask
// columns are set and cleared, so a row costs O(mask + streamed entries).
template <int MODE>
__global__ __launch_bounds__(64) void k_dense_global(Prob P, int cls, int32_t *__restrict__ gmap,
                                                     int32_t *__restrict__ gacc, int64_t ncols,
                                                     int32_t maxmask) {
    __shared__ int32_t pre[kMaxTaskA + 1];
    __shared__ int32_t bst[kMaxTaskA];
    __shared__ int32_t acnt[MODE == MODE_SUPPORT ? kMaxTaskA : 1];
    const int lane = threadIdx.x;
    int32_t *map = gmap + (size_t)blockIdx.x * ncols;
    int32_t *acc = gacc + (size_t)blockIdx.x * maxmask;
    const int32_t lo = P.bin_off[cls * kFlopBins], hi = P.bin_off[(cls + 1) * kFlopBins];
    uint32_t cnt = 0;
    for (int32_t ti = lo + blockIdx.x; ti < hi; ti += gridDim.x) {
        const Task tk = P.tasks[ti];
        const long long tstart = P.task_cycles ? clock64() : 0;
        const int32_t mb = P.mrst[tk.row], mlen = P.mrst[tk.row + 1] - mb;
        for (int32_t q = lane; q < mlen; q += 64) {
            map[P.mcol[mb + q]] = q + 1;
            if (HasAcc<MODE>::v) acc[q] = 0;
        }
        if (MODE == MODE_SUPPORT)
            for (int32_t x = lane; x < tk.a1 - tk.a0; x += 64) acnt[x] = 0;
        const int32_t total = build_prefix<64>(P, tk.a0, tk.a1, pre, bst, lane);
        __threadfence_block();
        __syncthreads();
        const uint32_t c = stream_entries<MODE, 64, 4>(P, tk, mb, pre, bst, total, lane, acc, acnt,
                                                       [&](int32_t b) -> int32_t { return map[b] - 1; });
        __threadfence_block();
        __syncthreads();
        flush<MODE, 64>(P, tk, mb, mlen, acc, acnt, lane);
        for (int32_t q = lane; q < mlen; q += 64) map[P.mcol[mb + q]] = 0;
        if (MODE == MODE_TC && P.rowcount) {
            uint32_t r = dev::wave_sum(c);
            if (lane == 0 && r) atomicAdd(&P.rowcount[tk.row], (unsigned long long)r);
        }
        cnt += c;
        if (P.task_cycles && threadIdx.x == 0) P.task_cycles[ti] = clock64() - tstart;
        __threadfence_block();
        __syncthreads();
    }
    finish_kernel<MODE>(P, cnt, cls, lane);
}

// Bitmaps of the hub rows of B: every nonempty row k < hub_n gets words
// [0, ceil(k / 32)) of its bitmap, built in LDS by one workgroup.  B is
// strictly lower, so every column of row k is below k.
__global__ __launch_bounds__(256) void k_hub_build(int32_t hub_n, int32_t hub_w, const int32_t *__restrict__ brst,
                                                   const int32_t *__restrict__ bcol, uint32_t *__restrict__ hub) {
    __shared__ uint32_t words[kWindow / 32];
    for (int32_t k = blockIdx.x; k < hub_n; k += gridDim.x) {
        const int32_t b0 = brst[k], b1 = brst[k + 1];
        if (b1 == b0) continue;
        const int32_t wk = (k + 31) >> 5;
        for (int32_t w = threadIdx.x; w < wk; w += blockDim.x) words[w] = 0;
        __syncthreads();
        for (int32_t p = b0 + threadIdx.x; p < b1; p += blockDim.x) {
            const int32_t c = bcol[p];
            atomicOr(&words[c >> 5], 1u << (c & 31));
        }
        __syncthreads();
        for (int32_t w = threadIdx.x; w < wk; w += blockDim.x) hub[(int64_t)k * hub_w + w] = words[w];
        __syncthreads();
    }
}

__global__ void k_zero_u32(uint32_t *p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// ---------------------------------------------------------------------------
// host driver

static int grid_for(gk_ctx *ctx, int64_t work, int per_cu) {
    int64_t g = (int64_t)ctx->num_cus * per_cu;
    if (work < g) g = work;
    return (int)(g < 1 ? 1 : g);
}

int zero_u32(gk_ctx *ctx, void *p, int64_t n) {
    if (n <= 0) return GK_OK;
    int blocks = grid_for(ctx, (n + 255) / 256, 16);
    k_zero_u32<<<blocks, 256, 0, ctx->stream>>>((uint32_t *)p, n);
    return launch_check("k_zero_u32");
}

int32_t hub_rows(const gk_dcsr *B, int mode, const gk_opts &o) {
    if (mode != MODE_TC || o.hub_window <= 0 || !B->lower || !B->sorted) return 0;
    return (int32_t)std::min<int64_t>(std::min<int64_t>(o.hub_window, kWindow), B->m);
}

int build_hub(gk_ctx *ctx, const gk_dcsr *B, int32_t hub_n, Prob *P) {
    P->hub = nullptr;
    P->hub_n = P->hub_w = 0;
    if (hub_n <= 0) return GK_OK;
    const int32_t hub_w = (hub_n + 31) >> 5;
    void *p;
    GK_TRY(scratch(ctx, SC_HUB, (size_t)hub_n * hub_w * 4, &p));
    const int blocks = grid_for(ctx, hub_n, 8);
    k_hub_build<<<blocks, 256, 0, ctx->stream>>>(hub_n, hub_w, B->rst, B->col, (uint32_t *)p);
    GK_TRY(launch_check("k_hub_build"));
    P->hub = (const uint32_t *)p;
    P->hub_n = hub_n;
    P->hub_w = hub_w;
    return GK_OK;
}

// Features only (for the selector's training and inference).
int row_stats(gk_ctx *ctx, const gk_dcsr *A, const gk_dcsr *B, const gk_dcsr *M, RowStats *st, int32_t hub_n,
              int32_t probe) {
    const int64_t m = A->m;
    void *p;
    size_t need = (size_t)m * (8 + 8 + 8 + 4 + 4 + 4 + 4 + 1) + 64;
    GK_TRY(scratch(ctx, SC_FEAT, need, &p));
    char *c = (char *)p;
    st->flop = (int64_t *)c; c += 8 * (size_t)m;
    st->sflop = (int64_t *)c; c += 8 * (size_t)m;
    st->mflop = (int64_t *)c; c += 8 * (size_t)m;
    st->mlen = (int32_t *)c; c += 4 * (size_t)m;
    st->alen = (int32_t *)c; c += 4 * (size_t)m;
    st->bmax = (int32_t *)c; c += 4 * (size_t)m;
    st->ntask = (int32_t *)c; c += 4 * (size_t)m;
    st->acc = (int8_t *)c;
    if (m == 0) return GK_OK;
    constexpr int G = 8;
    int blocks = grid_for(ctx, (m + 256 / G - 1) / (256 / G), 16);
    k_row_stats<G><<<blocks, 256, 0, ctx->stream>>>(m, A->rst, A->col, B->rst, M->rst, *st, hub_n,
                                                    A->rst == M->rst && A->col == M->col, probe);
    return launch_check("k_row_stats");
}

int plan_tasks(gk_ctx *ctx, const gk_dcsr *A, RowStats &st, const SelectParams &sp, Plan *plan) {
    const int64_t m = A->m;
    void *p;
    // Layout: u64 stats[GK_NUM_ACC * ST_NUM], u64 hits[GK_NUM_ACC], then the
    // int32 bin counters, offsets, cursors, per-accumulator max mask and the
    // class kernels' task tickets.
    const size_t u64s = kNumClass * ST_NUM + kNumClass;
    const size_t words = 3 * kNumBins + 1 + 2 * kNumClass;
    const size_t bytes = 8 * u64s + sizeof(int32_t) * words;
    GK_TRY(scratch(ctx, SC_BINS, bytes, &p));
    plan->stats = (unsigned long long *)p;
    plan->hits = plan->stats + kNumClass * ST_NUM;
    int32_t *bin_count = (int32_t *)(plan->hits + kNumClass), *bin_off = bin_count + kNumBins;
    int32_t *bin_cur = bin_off + kNumBins + 1;
    int32_t *acc_maxmask = bin_cur + kNumBins;
    plan->ticket = acc_maxmask + kNumClass;
    GK_HIP(hipMemsetAsync(p, 0, bytes, ctx->stream));
    int blocks = grid_for(ctx, (m + 255) / 256, 8);
    if (m > 0) {
        k_select_bin<<<blocks, 256, 0, ctx->stream>>>(m, st, sp, bin_count, acc_maxmask, plan->stats);
        GK_TRY(launch_check("k_select_bin"));
    }
    k_bin_scan<<<1, 64, 0, ctx->stream>>>(bin_count, bin_off, bin_cur);
    GK_TRY(launch_check("k_bin_scan"));
    // Task count and per-accumulator mask sizes come back to the host: the
    // hash-global and dense-global kernels size their scratch from them.
    int32_t *h;
    GK_TRY(pinned(ctx, sizeof(int32_t) * (kNumBins + 1 + kNumClass), (void **)&h));
    GK_HIP(hipMemcpyAsync(h, bin_off, sizeof(int32_t) * (kNumBins + 1), hipMemcpyDeviceToHost, ctx->stream));
    GK_HIP(hipMemcpyAsync(h + kNumBins + 1, acc_maxmask, sizeof(int32_t) * kNumClass, hipMemcpyDeviceToHost,
                          ctx->stream));
    GK_HIP(hipStreamSynchronize(ctx->stream));
    plan->ntasks = h[kNumBins];
    for (int a = 0; a < kNumClass; ++a) {
        plan->acc_tasks[a] = h[(a + 1) * kFlopBins] - h[a * kFlopBins];
        plan->acc_maxmask[a] = h[kNumBins + 1 + a];
    }
    void *tp;
    GK_TRY(scratch(ctx, SC_TASK, sizeof(Task) * (plan->ntasks + 1), &tp));
    plan->tasks = (Task *)tp;
    plan->bin_off = bin_off;
    if (m > 0) {
        k_scatter_tasks<<<blocks, 256, 0, ctx->stream>>>(m, st, sp, A->rst, bin_cur, plan->tasks);
        GK_TRY(launch_check("k_scatter_tasks"));
    }
    return GK_OK;
}

// Column-indexed maps for dense-global: `blocks` maps of `ncols` entries,
// kept all-zero between launches (each task clears what it set).
static int dense_maps(gk_ctx *ctx, int64_t blocks, int64_t ncols, int32_t **out) {
    if (ctx->dense_map && ctx->dense_blocks >= blocks && ctx->dense_ncols == ncols) {
        *out = ctx->dense_map;
        return GK_OK;
    }
    if (ctx->dense_map) {
        GK_HIP(hipStreamSynchronize(ctx->stream));
        GK_HIP(hipFree(ctx->dense_map));
        ctx->dense_map = nullptr;
    }
    size_t bytes = (size_t)blocks * (size_t)ncols * 4;
    if (hipMalloc(&ctx->dense_map, bytes ? bytes : 4) != hipSuccess) {
        ctx->dense_map = nullptr;
        set_error("hipMalloc(%zu) for dense-global maps failed", bytes);
        return GK_ENOMEM;
    }
    ctx->dense_blocks = blocks;
    ctx->dense_ncols = ncols;
    GK_TRY(zero_u32(ctx, ctx->dense_map, blocks * ncols));
    *out = ctx->dense_map;
    return GK_OK;
}

// Scratch of the global accumulators and the grid of every class, sized
// on the context stream before any launch (launch_classes itself only
// launches, so it can be captured in a hipGraph).
int setup_classes(gk_ctx *ctx, const Plan &plan, int64_t bn, const gk_opts &o, ClassLaunch *L) {
    auto per_cu = [](int32_t v, int dflt) { return v > 0 ? std::min<int32_t>(v, 64) : dflt; };
    const int b_small = per_cu(o.blocks_small, kLdsSmallBlocksPerCu), b_long = per_cu(o.blocks_long, kWindowBlocksLong);
    const int b_mca = per_cu(o.blocks_mca, kMcaBlocksPerCu), b_glob = per_cu(o.blocks_global, kGlobalBlocksPerCu);
    *L = ClassLaunch{};
    L->bn = bn;
    L->grid[kClassMsaLarge] = plan.acc_tasks[kClassMsaLarge] ? grid_for(ctx, plan.acc_tasks[kClassMsaLarge], b_long) : 0;
    L->grid_lds_large = plan.acc_tasks[kClassMsaLarge]
                            ? grid_for(ctx, plan.acc_tasks[kClassMsaLarge], std::min(b_long, kLdsBlocksPerCu))
                            : 0;
    L->grid[GK_ACC_MSA_LDS] = plan.acc_tasks[GK_ACC_MSA_LDS] ? grid_for(ctx, plan.acc_tasks[GK_ACC_MSA_LDS], b_small) : 0;
    L->grid[GK_ACC_MCA_WAVE] =
        plan.acc_tasks[GK_ACC_MCA_WAVE] ? grid_for(ctx, (plan.acc_tasks[GK_ACC_MCA_WAVE] + 3) / 4, b_mca) : 0;
    if (plan.acc_tasks[GK_ACC_HASH_GLOBAL]) {
        while ((1 << L->hlg) < 2 * plan.acc_maxmask[GK_ACC_HASH_GLOBAL]) ++L->hlg;
        int hg = grid_for(ctx, plan.acc_tasks[GK_ACC_HASH_GLOBAL], b_glob);
        size_t cap = (size_t)1 << L->hlg;
        // Bound the tables to about 4 GiB.
        while (hg > 1 && (size_t)hg * cap * 10 > ((size_t)4 << 30)) hg >>= 1;
        void *sp;
        GK_TRY(scratch(ctx, SC_ACC, (size_t)hg * cap * 4 * 2 + (size_t)hg * (cap / 2) * 4, &sp));
        L->hkey = (int32_t *)sp;
        L->hpos = L->hkey + (size_t)hg * cap;
        L->hacc = L->hpos + (size_t)hg * cap;
        L->grid[GK_ACC_HASH_GLOBAL] = hg;
    }
    if (plan.acc_tasks[GK_ACC_DENSE_GLOBAL]) {
        int dg = grid_for(ctx, plan.acc_tasks[GK_ACC_DENSE_GLOBAL], b_glob);
        // Bound the map memory to about 4 GiB.
        while (dg > 1 && (size_t)dg * (size_t)bn * 4 > ((size_t)4 << 30)) dg >>= 1;
        L->dmm = plan.acc_maxmask[GK_ACC_DENSE_GLOBAL];
        GK_TRY(dense_maps(ctx, dg, bn, &L->dmap));
        void *sp;
        GK_TRY(scratch(ctx, SC_TMP7, (size_t)dg * (size_t)L->dmm * 4 + 4, &sp));
        L->dacc = (int32_t *)sp;
        L->grid[GK_ACC_DENSE_GLOBAL] = dg;
    }
    return GK_OK;
}

// Launch the accumulator kernels of one product.  Each class runs on its own
// auxiliary stream (forked from the context stream, joined back before the
// caller continues), so the classes share the GPU instead of running one
// after the other.  With tm, each class is bracketed by its timer.
template <int MODE>
int launch_classes(gk_ctx *ctx, const Prob &P, const ClassLaunch &L, Timer *tm) {
    static_assert(kNumClass <= gk_ctx::kAux + 1, "one stream per kernel class");
    hipStream_t sc[kNumClass];
    sc[GK_ACC_DENSE_GLOBAL] = ctx->stream;
    sc[kClassMsaLarge] = ctx->aux[0];
    sc[GK_ACC_MSA_LDS] = ctx->aux[1];
    sc[GK_ACC_MCA_WAVE] = ctx->aux[2];
    sc[GK_ACC_HASH_GLOBAL] = ctx->aux[3];
    // the class kernels' task tickets start at 0 (a fill kernel, not a
    // memset node, so that a captured graph replays it correctly)
    k_zero_u32<<<1, 64, 0, ctx->stream>>>((uint32_t *)P.ticket, kNumClass);
    GK_TRY(launch_check("k_zero_u32"));
    GK_HIP(hipEventRecord(ctx->fork, ctx->stream));
    for (int i = 0; i < gk_ctx::kAux; ++i) GK_HIP(hipStreamWaitEvent(ctx->aux[i], ctx->fork, 0));
    if (L.grid[kClassMsaLarge]) {
        if (tm) tm[kClassMsaLarge].start();
        constexpr int WM = MODE == MODE_VALUES ? MODE_TC : MODE;
        if (MODE == MODE_TC && P.hub)
            k_msa_window<MODE_TC, kWindowWavesLong, kWindowHashLong, kLdsMaxMask, true>
                <<<L.grid[kClassMsaLarge], 64 * kWindowWavesLong, 0, sc[kClassMsaLarge]>>>(P, kClassMsaLarge);
        else if (MODE == MODE_TC || MODE == MODE_SUPPORT)
            k_msa_window<WM, kWindowWavesLong, kWindowHashLong, kLdsMaxMask, false>
                <<<L.grid[kClassMsaLarge], 64 * kWindowWavesLong, 0, sc[kClassMsaLarge]>>>(P, kClassMsaLarge);
        else
            k_msa_lds<MODE, kLdsHashCap, kLdsMaxMask><<<L.grid_lds_large, 64, 0, sc[kClassMsaLarge]>>>(P,
                                                                                                       kClassMsaLarge);
        if (tm) tm[kClassMsaLarge].stop();
        GK_TRY(launch_check("k_msa_lds<large>"));
    }
    if (L.grid[GK_ACC_MSA_LDS]) {
        if (tm) tm[GK_ACC_MSA_LDS].start();
        constexpr int WM = MODE == MODE_VALUES ? MODE_TC : MODE;
        if (MODE == MODE_TC && P.hub)
            k_msa_window<MODE_TC, 1, kLdsSmallCap, kLdsSmallMask, true>
                <<<L.grid[GK_ACC_MSA_LDS], 64, 0, sc[GK_ACC_MSA_LDS]>>>(P, GK_ACC_MSA_LDS);
        else if (MODE == MODE_TC || MODE == MODE_SUPPORT)
            k_msa_window<WM, 1, kLdsSmallCap, kLdsSmallMask, false>
                <<<L.grid[GK_ACC_MSA_LDS], 64, 0, sc[GK_ACC_MSA_LDS]>>>(P, GK_ACC_MSA_LDS);
        else
            k_msa_lds<MODE, kLdsSmallCap, kLdsSmallMask><<<L.grid[GK_ACC_MSA_LDS], 64, 0, sc[GK_ACC_MSA_LDS]>>>(
                P, GK_ACC_MSA_LDS);
        if (tm) tm[GK_ACC_MSA_LDS].stop();
        GK_TRY(launch_check("k_msa_lds<small>"));
    }
    if (L.grid[GK_ACC_MCA_WAVE]) {
        if (tm) tm[GK_ACC_MCA_WAVE].start();
        if (MODE == MODE_TC && P.hub)
            k_mca_wave<MODE_TC, 16, true><<<L.grid[GK_ACC_MCA_WAVE], 64, 0, sc[GK_ACC_MCA_WAVE]>>>(P, GK_ACC_MCA_WAVE);
        else
            k_mca_wave<MODE, 16, false><<<L.grid[GK_ACC_MCA_WAVE], 64, 0, sc[GK_ACC_MCA_WAVE]>>>(P, GK_ACC_MCA_WAVE);
        if (tm) tm[GK_ACC_MCA_WAVE].stop();
        GK_TRY(launch_check("k_mca_wave"));
    }
    if (L.grid[GK_ACC_HASH_GLOBAL]) {
        if (tm) tm[GK_ACC_HASH_GLOBAL].start();
        k_hash_global<MODE><<<L.grid[GK_ACC_HASH_GLOBAL], 64, 0, sc[GK_ACC_HASH_GLOBAL]>>>(
            P, GK_ACC_HASH_GLOBAL, L.hkey, L.hpos, L.hacc, L.hlg);
        if (tm) tm[GK_ACC_HASH_GLOBAL].stop();
        GK_TRY(launch_check("k_hash_global"));
    }
    if (L.grid[GK_ACC_DENSE_GLOBAL]) {
        if (tm) tm[GK_ACC_DENSE_GLOBAL].start();
        k_dense_global<MODE><<<L.grid[GK_ACC_DENSE_GLOBAL], 64, 0, sc[GK_ACC_DENSE_GLOBAL]>>>(
            P, GK_ACC_DENSE_GLOBAL, L.dmap, L.dacc, L.bn, L.dmm);
        if (tm) tm[GK_ACC_DENSE_GLOBAL].stop();
        GK_TRY(launch_check("k_dense_global"));
    }
    for (int i = 0; i < gk_ctx::kAux; ++i) {
        GK_HIP(hipEventRecord(ctx->join[i], ctx->aux[i]));
        GK_HIP(hipStreamWaitEvent(ctx->stream, ctx->join[i], 0));
    }
    return GK_OK;
}
template int launch_classes<MODE_SUPPORT>(gk_ctx *, const Prob &, const ClassLaunch &, Timer *);

template <int MODE>
static int run_mode(gk_ctx *ctx, const Prob &P, const Plan &plan, int64_t bn, const gk_opts &o, double *acc_ms) {
    ClassLaunch L;
    GK_TRY(setup_classes(ctx, plan, bn, o, &L));
    hipStream_t sc[kNumClass] = {ctx->aux[1], ctx->aux[2], ctx->aux[3], ctx->stream, ctx->aux[0]};
    Timer tm[kNumClass] = {Timer(ctx, sc[0]), Timer(ctx, sc[1]), Timer(ctx, sc[2]), Timer(ctx, sc[3]),
                           Timer(ctx, sc[4])};
    GK_TRY(launch_classes<MODE>(ctx, P, L, tm));
    for (int a = 0; a < kNumClass; ++a) acc_ms[a] = L.grid[a] ? tm[a].ms() : 0.0;
    return GK_OK;
}

gk_opts mode_opts(const gk_opts &o, int mode) {
    const bool tc = mode == MODE_TC;
    gk_opts r = o;
    if (r.mca_max_mask < 0) r.mca_max_mask = tc ? 48 : 16;
    if (r.hub_probe < 0) r.hub_probe = 2;
    if (r.blocks_long <= 0) r.blocks_long = tc ? 16 : kWindowBlocksLong;
    if (r.blocks_mca <= 0) r.blocks_mca = tc ? 16 : kMcaBlocksPerCu;
    return r;
}

SelectParams select_params(const gk_opts &o) {
    SelectParams sp;
    sp.selector = o.selector;
    sp.forced = o.forced_accum;
    sp.mca_max_mask = o.mca_max_mask;
    sp.lds_max_mask = o.lds_max_mask;
    sp.dense_min_mask = o.dense_min_mask;
    sp.split_flop = o.split_flop > 0 ? o.split_flop : (1 << 18);
    sp.small_max_mask = std::max(1, std::min<int32_t>(o.small_max_mask > 0 ? o.small_max_mask : kLdsSmallMask,
                                                      kLdsSmallMask));
    sp.hub = 0;
    sp.margin = o.learned_margin / 1000.0f;
    return sp;
}

int masked_product(gk_ctx *ctx, const gk_dcsr *A, const gk_dcsr *B, const gk_dcsr *M, int mode, int32_t *d_out,
                   unsigned long long *d_count, unsigned long long *d_rowcount, const gk_opts &o_in,
                   ProductInfo *info, long long *task_cycles, Plan *plan_out) {
    const gk_opts o = mode_opts(o_in, mode);
    if (A->n != B->m || A->m != M->m || B->n != M->n) {
        set_error("shape mismatch: A %lldx%lld, B %lldx%lld, M %lldx%lld", (long long)A->m, (long long)A->n,
                  (long long)B->m, (long long)B->n, (long long)M->m, (long long)M->n);
        return GK_EINVAL;
    }
    Timer total(ctx), sel(ctx);
    total.start();
    sel.start();
    RowStats st;
    const int32_t hub_n = hub_rows(B, mode, o);
    GK_TRY(row_stats(ctx, A, B, M, &st, hub_n, std::max(o.hub_probe, 0)));
    SelectParams sp = select_params(o);
    sp.hub = hub_n > 0;
    if (sp.selector == GK_SEL_LEARNED)
        GK_TRY(model_select(ctx, st, A->m, A->rst == M->rst && A->col == M->col, sp));
    else if (sp.selector == kSelMixed)
        GK_TRY(mixed_assign(ctx, st, A->m, o.forced_accum));
    Plan plan;
    GK_TRY(plan_tasks(ctx, A, st, sp, &plan));
    Prob P{};
    GK_TRY(build_hub(ctx, B, hub_n, &P));
    P.probe = (P.hub && A->rst == M->rst && A->col == M->col) ? std::max(o.hub_probe, 0) : 0;
    sel.stop();
    if (mode == MODE_VALUES || mode == MODE_SUPPORT) GK_TRY(zero_u32(ctx, d_out, M->nz));
    P.arst = A->rst; P.acol = A->col; P.aval = A->val;
    P.brst = B->rst; P.bcol = B->col; P.bval = B->val;
    P.mrst = M->rst; P.mcol = M->col;
    P.cval = d_out; P.count = d_count; P.rowcount = d_rowcount; P.hits = plan.hits;
    P.task_cycles = task_cycles;
    P.tasks = plan.tasks; P.bin_off = plan.bin_off; P.ticket = plan.ticket;
    double acc_ms[kNumClass] = {};
    if (mode == MODE_TC) GK_TRY(run_mode<MODE_TC>(ctx, P, plan, B->n, o, acc_ms));
    else if (mode == MODE_VALUES) GK_TRY(run_mode<MODE_VALUES>(ctx, P, plan, B->n, o, acc_ms));
    else GK_TRY(run_mode<MODE_SUPPORT>(ctx, P, plan, B->n, o, acc_ms));
    total.stop();
    if (plan_out) *plan_out = plan;
    if (info) {
        unsigned long long hs[kNumClass * (ST_NUM + 1)];
        GK_HIP(d2h(ctx, hs, plan.stats, sizeof hs));
        const bool same_am = A->col == M->col && A->rst == M->rst;
        *info = ProductInfo{};
        info->total_ms = total.ms();
        info->select_ms = sel.ms();
        info->ntasks = plan.ntasks;
        for (int c = 0; c < kNumClass; ++c) {
            const int a = c == kClassMsaLarge ? (int)GK_ACC_MSA_LDS : c;  // reported accumulator
            const unsigned long long *x = hs + c * ST_NUM;
            const double hits = (double)hs[kNumClass * ST_NUM + c];
            // Algorithmic HBM bytes: mask row pointers and columns, A columns
            // unless A is the mask, a B row-pointer pair per A entry, the
            // streamed B columns, and the outputs (C values once; supports:
            // three 4-byte increments per triangle).
            double b = 8.0 * x[ST_TASKS] + 4.0 * x[ST_MASK] + (same_am ? 0.0 : 4.0 * x[ST_AROW]) +
                       8.0 * x[ST_AROW] + 4.0 * x[ST_FLOP];
            if (mode == MODE_VALUES) b += 4.0 * x[ST_MASK];
            if (mode == MODE_SUPPORT) b += 12.0 * hits;
            info->acc_tasks[a] += plan.acc_tasks[c];
            info->class_tasks[c] = plan.acc_tasks[c];
            info->acc_ms[c] = acc_ms[c];
            info->acc_bytes[c] = b;
            info->kernel_ms += acc_ms[c];
            info->flop += (double)x[ST_FLOP];
            info->hits += hits;
        }
    }
    return GK_OK;
}

int prepare_product(gk_ctx *ctx, const gk_dcsr *A, const gk_dcsr *B, const gk_dcsr *M, int32_t *d_out,
                    const gk_opts &o_in, Prob *P, ClassLaunch *L) {
    const gk_opts o = mode_opts(o_in, MODE_SUPPORT);
    if (A->n != B->m || A->m != M->m || B->n != M->n) {
        set_error("shape mismatch in a prepared product");
        return GK_EINVAL;
    }
    RowStats st;
    GK_TRY(row_stats(ctx, A, B, M, &st));
    SelectParams sp = select_params(o);
    if (sp.selector == GK_SEL_LEARNED)
        GK_TRY(model_select(ctx, st, A->m, A->rst == M->rst && A->col == M->col, sp));
    Plan plan;
    GK_TRY(plan_tasks(ctx, A, st, sp, &plan));
    *P = Prob{};
    P->arst = A->rst; P->acol = A->col; P->aval = A->val;
    P->brst = B->rst; P->bcol = B->col; P->bval = B->val;
    P->mrst = M->rst; P->mcol = M->col;
    P->cval = d_out; P->hits = plan.hits;
    P->tasks = plan.tasks; P->bin_off = plan.bin_off; P->ticket = plan.ticket;
    return setup_classes(ctx, plan, B->n, o, L);
}

void record_timing(gk_ctx *ctx, const ProductInfo &pi, bool reset) {
    gk_timing &t = ctx->last;
    if (reset) t = gk_timing{};
    t.total_ms += pi.total_ms;
    t.kernel_ms += pi.kernel_ms;
    t.select_ms += pi.select_ms;
    t.launches += 1;
    for (int a = 0; a < GK_NUM_ACC; ++a) t.rows_per_acc[a] += pi.acc_tasks[a];
    for (int c = 0; c < kNumClass; ++c) {
        t.acc_ms[c] += pi.acc_ms[c];
        t.acc_bytes[c] += pi.acc_bytes[c];
        t.acc_launches[c] += pi.class_tasks[c] ? 1 : 0;
        t.bytes_alg += pi.acc_bytes[c];
    }
    t.flop += pi.flop;
    t.hits += pi.hits;
}

}  // namespace gk

// ---------------------------------------------------------------------------
// C ABI

extern "C" int gk_masked_spgemm_dev(gk_ctx *ctx, const gk_dcsr *A, const gk_dcsr *B, const gk_dcsr *M,
                                    int32_t *d_cval, const gk_opts *opts) {
    GK_CHECK_ARG(ctx && A && B && M && (d_cval || M->nz == 0), "null argument");
    GK_TRY(require_sorted(M, "M"));
    gk_opts o = resolve_opts(opts);
    ProductInfo pi;
    GK_TRY(masked_product(ctx, A, B, M, MODE_VALUES, d_cval, nullptr, nullptr, o, &pi));
    record_timing(ctx, pi, true);
    return GK_OK;
}

extern "C" int gk_masked_spgemm_i32(gk_ctx *ctx, const gk_csr *A, const gk_csr *B, const gk_csr *M, int32_t *c_val,
                                    const gk_opts *opts) {
    GK_CHECK_ARG(ctx && M && (c_val || M->nz == 0), "null argument");
    GK_TRY(check_csr(A, "A"));
    GK_TRY(check_csr(B, "B"));
    GK_TRY(check_csr(M, "M"));
    gk_dcsr dA, dB, dM;
    GK_TRY(upload_csr(ctx, A, &dA));
    int rc = upload_csr(ctx, B, &dB);
    if (rc) { free_csr(&dA); return rc; }
    rc = upload_csr(ctx, M, &dM);
    if (rc) { free_csr(&dA); free_csr(&dB); return rc; }
    int32_t *d_c = nullptr;
    if (hipMalloc(&d_c, sizeof(int32_t) * (M->nz ? M->nz : 1)) != hipSuccess) {
        free_csr(&dA); free_csr(&dB); free_csr(&dM);
        set_error("hipMalloc for C failed");
        return GK_ENOMEM;
    }
    rc = gk_masked_spgemm_dev(ctx, &dA, &dB, &dM, d_c, opts);
    if (rc == GK_OK && M->nz) {
        if (d2h(ctx, c_val, d_c, sizeof(int32_t) * M->nz) != hipSuccess) {
            set_error("copy of C to host failed");
            rc = GK_EHIP;
        }
    }
    (void)hipFree(d_c);
    free_csr(&dA); free_csr(&dB); free_csr(&dM);
    return rc;
}

extern "C" int gk_tc_count_dev(gk_ctx *ctx, const gk_dcsr *L, uint64_t *count, const gk_opts *opts) {
    GK_CHECK_ARG(ctx && L && count, "null argument");
    GK_CHECK_ARG(L->m == L->n, "L must be square");
    GK_TRY(require_sorted(L, "L"));
    gk_opts o = resolve_opts(opts);
    void *p;
    GK_TRY(scratch(ctx, SC_RED, 64, &p));
    unsigned long long *d_count = (unsigned long long *)p;
    GK_HIP(hipMemsetAsync(d_count, 0, 8, ctx->stream));
    ProductInfo pi;
    GK_TRY(masked_product(ctx, L, L, L, MODE_TC, nullptr, d_count, nullptr, o, &pi));
    unsigned long long *h;
    GK_TRY(pinned(ctx, 8, (void **)&h));
    GK_HIP(hipMemcpyAsync(h, d_count, 8, hipMemcpyDeviceToHost, ctx->stream));
    GK_HIP(hipStreamSynchronize(ctx->stream));
    *count = *h;
    record_timing(ctx, pi, true);
    return GK_OK;
}

extern "C" int gk_tc_rows_dev(gk_ctx *ctx, const gk_dcsr *L, uint64_t *d_rowcount, const gk_opts *opts) {
    GK_CHECK_ARG(ctx && L && d_rowcount, "null argument");
    GK_CHECK_ARG(L->m == L->n, "L must be square");
    GK_TRY(require_sorted(L, "L"));
    gk_opts o = resolve_opts(opts);
    void *p;
    GK_TRY(scratch(ctx, SC_RED, 64, &p));
    unsigned long long *d_count = (unsigned long long *)p;
    GK_HIP(hipMemsetAsync(d_count, 0, 8, ctx->stream));
    GK_HIP(hipMemsetAsync(d_rowcount, 0, 8 * (size_t)L->m, ctx->stream));
    ProductInfo pi;
    GK_TRY(masked_product(ctx, L, L, L, MODE_TC, nullptr, d_count, (unsigned long long *)d_rowcount, o, &pi));
    record_timing(ctx, pi, true);
    return GK_OK;
}

extern "C" int gk_tc_count(gk_ctx *ctx, const gk_csr *L, uint64_t *count, const gk_opts *opts) {
    GK_CHECK_ARG(ctx && count, "null argument");
    gk_dcsr d;
    GK_TRY(upload_csr(ctx, L, &d));
    int rc = gk_tc_count_dev(ctx, &d, count, opts);
    free_csr(&d);
    return rc;
}
