// Batching: the triangle count and the k-truss of many independent graphs
// in one launch sequence, captured once as a hipGraph and replayed.
//
// The graphs arrive as one block-diagonal symmetric pattern A (graph g on
// vertices [voff[g], voff[g+1])), resident in HBM.  Graph g's edges are
// then the contiguous range [eoff[g], eoff[g+1]) of the union's strictly
// lower part L.  One step:
//   1. supports of every edge: one SUPPORT-mode masked product C<L> = L*L
//      over the union -- the engine's persistent class kernels balance
//      rows across graphs whatever their sizes;
//   2. k_peel_graphs: a persistent kernel; each workgroup takes graphs from
//      a queue (largest first), sums the graph's supports as it collects
//      its first frontier (triangles(g) = that sum / 3: each triangle adds
//      one to each of its three edges) and peels the graph to the k-truss
//      inside the workgroup, rounds separated by workgroup barriers -- the
//      rounds of src/main.cpp:102-133 for every graph, without a grid-wide
//      sync;
//   3. k_keep_flags: per entry of A whether it is in its graph's k-truss.
// Every buffer the step writes is cleared by kernels of the step itself,
// so replays of the captured graph repeat the first run exactly.
#include <algorithm>
#include <chrono>
#include <numeric>
#include <vector>

#include "gk_device.h"
#include "spgemm.h"
#include "truss.h"

using namespace gk;

namespace gk {
namespace {

constexpr int kPeelThreads = 1024;
constexpr int kBatchMcaBlocksPerCu = 32;
// consecutive streamed items per thread and step: four keep the peeling
// kernel at 64 VGPRs, so two 1024-thread workgroups share a CU (eight waves
// per SIMD); eight items took 124 VGPRs and one workgroup per CU (4096-graph
// step 91.8 -> 80.1 ms)
constexpr int kPeelItems = 4;

template <int NT>
__device__ __forceinline__ int32_t block_incl_scan(int32_t v, int32_t *lds, int32_t *total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int32_t incl = dev::wave_incl_scan(v);
    if (lane == 63) lds[w] = incl;
    __syncthreads();
    int32_t before = 0, tot = 0;
#pragma unroll
    for (int x = 0; x < NT / 64; ++x) {
        const int32_t t = lds[x];
        before += x < w ? t : 0;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return before + incl;
}

__global__ void k_zero_bytes(int64_t n, uint8_t *__restrict__ p) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// One workgroup peels one graph at a time: frontier F = alive edges with
// support < thresh; per round each frontier edge f = (y, x) streams the row
// of x (y = endpoint with the longer row, searched by bisection); a
// triangle (f, a = (y,w), b = (x,w)) is destroyed by its frontier edge with
// the smallest id, which takes one support off each of its alive edges;
// an edge whose support drops below thresh joins the next frontier.
// States: 0 alive, 1 frontier, 2 removed.  Graph g uses the segments
// [eoff[g], eoff[g+1]) of buf0 / buf1 / fy / fx and eoff[g] + g of pre.
__global__ __launch_bounds__(kPeelThreads) __attribute__((amdgpu_waves_per_eu(8))) void k_peel_graphs(
    int32_t ngraphs, const int32_t *__restrict__ order, const int32_t *__restrict__ eoff,
    const int32_t *__restrict__ arst, const int32_t *__restrict__ acol, const int32_t *__restrict__ aeid,
    const int32_t *__restrict__ eu, const int32_t *__restrict__ ev, int32_t *__restrict__ sup,
    uint8_t *__restrict__ st, int32_t thresh, int32_t *__restrict__ buf0, int32_t *__restrict__ buf1,
    int32_t *__restrict__ fy, int32_t *__restrict__ fx, int32_t *__restrict__ fh, int32_t *__restrict__ pre,
    int32_t *__restrict__ queue, long long *__restrict__ per_graph) {
    // per_graph[g]: the edges of g's k-truss; per_graph[ngraphs + g]: g's triangles
    long long *alive_edges = per_graph, *tri = per_graph + ngraphs;
    __shared__ int32_t lds[kPeelThreads / 64 + 1];
    __shared__ unsigned long long s_sum[kPeelThreads / 64];
    __shared__ int32_t s_g, s_n;
    __shared__ int32_t s_alive[kPeelThreads / 64];
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) s_g = atomicAdd(queue, 1);
        __syncthreads();
        const int32_t gi = s_g;
        __syncthreads();
        if (gi >= ngraphs) return;
        const int32_t g = order[gi];
        const int32_t e0 = eoff[g], e1 = eoff[g + 1];
        int32_t *F = buf0 + e0, *Fn = buf1 + e0;
        // per frontier edge f: y's row [gy[f], gh[f]) and gx[f] = the start of
        // x's row minus the items before f (item it of f is entry gx[f] + it)
        int32_t *gy = fy + e0, *gx = fx + e0, *gh = fh + e0, *gp = pre + e0 + g;
        if (tid == 0) s_n = 0;
        __syncthreads();
        unsigned long long ssum = 0;  // the graph's supports, three per triangle
        for (int32_t base = e0; base < e1; base += kPeelThreads) {
            const int32_t e = base + tid;
            const int32_t sv = e < e1 ? sup[e] : 0;
            ssum += (uint32_t)sv;
            dev::block_append<kPeelThreads>(e < e1 && sv < thresh, e, F, &s_n, lds);
        }
        ssum = dev::wave_sum(ssum);
        if ((tid & 63) == 0) s_sum[tid >> 6] = ssum;
        __syncthreads();
        if (tid == 0) {
            unsigned long long t = 0;
#pragma unroll
            for (int x = 0; x < kPeelThreads / 64; ++x) t += s_sum[x];
            tri[g] = (long long)(t / 3);
        }
        int32_t nF = s_n;
        while (nF > 0) {
            // mark the frontier, pick the streamed endpoint, scan the lengths
            int32_t carry = 0;
            for (int32_t c = 0; c < nF; c += kPeelThreads) {
                const int32_t f = c + tid;
                int32_t len = 0, xs = 0;
                if (f < nF) {
                    const int32_t e = F[f], u = eu[e], v = ev[e];
                    st[e] = 1;
                    const int32_t ru = arst[u], ru1 = arst[u + 1], rv = arst[v], rv1 = arst[v + 1];
                    const int32_t du = ru1 - ru, dv = rv1 - rv;
                    const bool uy = du > dv || (du == dv && u > v);
                    gy[f] = uy ? ru : rv;
                    gh[f] = uy ? ru1 : rv1;
                    xs = uy ? rv : ru;
                    // an edge with no intact triangle has nothing to destroy
                    len = sup[e] > 0 ? (uy ? dv : du) : 0;
                }
                int32_t tot;
                const int32_t incl = block_incl_scan<kPeelThreads>(len, lds, &tot);
                if (f < nF) {
                    gp[f] = carry + incl - len;
                    gx[f] = xs - (carry + incl - len);
                }
                carry += tot;
            }
            if (tid == 0) {
                gp[nF] = carry;
                s_n = 0;
            }
            __syncthreads();
            const int32_t total = carry;
            // Each thread takes kPeelItems consecutive items per step: one
            // search over the frontier prefix finds the owner of its first
            // item, a cursor walks to the next owners; then every stage's
            // loads are issued for all the thread's items before any is
            // used: the entries (w, edge b) -> b's state -> the lower bounds
            // of w in y's sorted row, one bisection step of every item at a
            // time -> the edge a of a found w and its state -> the support
            // updates.
            for (int32_t base = 0; base < total; base += kPeelThreads * kPeelItems) {
                const int32_t it0 = base + tid * kPeelItems;
                bool xk[2 * kPeelItems];
                int32_t xv[2 * kPeelItems];
#pragma unroll
                for (int j = 0; j < 2 * kPeelItems; ++j) {
                    xk[j] = false;
                    xv[j] = 0;
                }
                int32_t pp[kPeelItems], ee[kPeelItems], lo[kPeelItems], hi[kPeelItems];
#pragma unroll
                for (int j = 0; j < kPeelItems; ++j) {
                    pp[j] = -1;
                    ee[j] = lo[j] = hi[j] = 0;
                }
                if (it0 < total) {
                    int32_t l = 0, h = nF - 1;
                    while (l < h) {  // largest f with gp[f] <= it0
                        const int32_t mid = l + ((h - l + 1) >> 1);
                        if (gp[mid] <= it0) l = mid; else h = mid - 1;
                    }
                    int32_t f = l, fend = gp[f + 1];
                    int32_t e = F[f], p0 = gx[f], yl = gy[f], yh = gh[f];
#pragma unroll
                    for (int j = 0; j < kPeelItems; ++j) {
                        const int32_t it = it0 + j;
                        if (it < total) {
                            if (it >= fend) {
                                while (it >= fend) fend = gp[++f + 1];
                                e = F[f];
                                p0 = gx[f];
                                yl = gy[f];
                                yh = gh[f];
                            }
                            pp[j] = p0 + it;
                            ee[j] = e;
                            lo[j] = yl;
                            hi[j] = yh;
                        }
                    }
                }
                int32_t w[kPeelItems], b[kPeelItems];
#pragma unroll
                for (int j = 0; j < kPeelItems; ++j) {
                    w[j] = pp[j] >= 0 ? acol[pp[j]] : 0;
                    b[j] = pp[j] >= 0 ? aeid[pp[j]] : 0;
                }
                uint8_t s2[kPeelItems];
#pragma unroll
                for (int j = 0; j < kPeelItems; ++j) s2[j] = pp[j] >= 0 ? st[b[j]] : (uint8_t)2;
                // items whose edge b is removed (or a frontier edge with a
                // smaller id, which owns the triangle) search nothing
#pragma unroll
                for (int j = 0; j < kPeelItems; ++j)
                    if (s2[j] == 2 || (s2[j] == 1 && b[j] < ee[j])) hi[j] = lo[j];
                int32_t pos[kPeelItems];
#pragma unroll
                for (int j = 0; j < kPeelItems; ++j) pos[j] = -1;
                for (;;) {  // lower bound of w in acol[lo, hi): rows hold each column once
                    bool any = false;
                    int32_t mid[kPeelItems], v[kPeelItems];
#pragma unroll
                    for (int j = 0; j < kPeelItems; ++j) {
                        mid[j] = lo[j] + ((hi[j] - lo[j]) >> 1);
                        v[j] = lo[j] < hi[j] ? acol[mid[j]] : 0;
                    }
#pragma unroll
                    for (int j = 0; j < kPeelItems; ++j) {
                        if (lo[j] >= hi[j]) continue;
                        if (v[j] == w[j]) {
                            pos[j] = mid[j];
                            hi[j] = lo[j];
                        } else if (v[j] < w[j]) {
                            lo[j] = mid[j] + 1;
                        } else {
                            hi[j] = mid[j];
                        }
                        any |= lo[j] < hi[j];
                    }
                    if (!any) break;
                }
                int32_t a[kPeelItems];
#pragma unroll
                for (int j = 0; j < kPeelItems; ++j) a[j] = pos[j] >= 0 ? aeid[pos[j]] : -1;
                uint8_t s1[kPeelItems];
#pragma unroll
                for (int j = 0; j < kPeelItems; ++j) s1[j] = a[j] >= 0 ? st[a[j]] : (uint8_t)2;
#pragma unroll
                for (int j = 0; j < kPeelItems; ++j) {
                    if (a[j] < 0 || s1[j] == 2 || (s1[j] == 1 && a[j] < ee[j])) continue;
                    if (s1[j] == 0 && atomicSub(&sup[a[j]], 1) == thresh) {
                        xk[2 * j] = true;
                        xv[2 * j] = a[j];
                    }
                    if (s2[j] == 0 && atomicSub(&sup[b[j]], 1) == thresh) {
                        xk[2 * j + 1] = true;
                        xv[2 * j + 1] = b[j];
                    }
                }
                dev::block_append_n<kPeelThreads, 2 * kPeelItems>(xk, xv, Fn, &s_n, lds);
            }
            for (int32_t f = tid; f < nF; f += kPeelThreads) st[F[f]] = 2;
            __syncthreads();
            nF = s_n;
            int32_t *t = F;
            F = Fn;
            Fn = t;
            __syncthreads();
        }
        // edges left in the graph's k-truss
        int32_t c = 0;
        for (int32_t e = e0 + tid; e < e1; e += kPeelThreads) c += st[e] == 0;
        c = dev::wave_sum(c);
        if ((tid & 63) == 0) s_alive[tid >> 6] = c;
        __syncthreads();
        if (tid == 0) {
            long long a = 0;
            for (int x = 0; x < kPeelThreads / 64; ++x) a += s_alive[x];
            alive_edges[g] = a;
        }
        __syncthreads();
    }
}

__global__ void k_keep_flags(int64_t nz, const int32_t *__restrict__ aeid, const uint8_t *__restrict__ st,
                             uint8_t *__restrict__ keep) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < nz; p += (int64_t)gridDim.x * blockDim.x)
        keep[p] = st[aeid[p]] == 0;
}

// Flags an entry of A whose column lies outside its row's graph.
__global__ void k_check_blocks(int64_t m, int64_t ngraphs, const int64_t *__restrict__ voff,
                               const int32_t *__restrict__ rst, const int32_t *__restrict__ col,
                               int32_t *__restrict__ bad) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        int64_t lo = 0, hi = ngraphs - 1;
        while (lo < hi) {  // largest g with voff[g] <= i
            const int64_t mid = lo + ((hi - lo + 1) >> 1);
            if (voff[mid] <= i) lo = mid; else hi = mid - 1;
        }
        const int64_t b = voff[lo], e = voff[lo + 1];
        for (int32_t p = rst[i]; p < rst[i + 1]; ++p)
            if (col[p] < b || col[p] >= e) *bad = 1;
    }
}

__global__ void k_gather_offsets(int64_t n, const int64_t *__restrict__ voff, const int32_t *__restrict__ lrst,
                                 int32_t *__restrict__ eoff) {
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n; g += (int64_t)gridDim.x * blockDim.x)
        eoff[g] = lrst[voff[g]];
}

struct DevMem {
    std::vector<void *> mem;
    ~DevMem() {
        for (void *p : mem) (void)hipFree(p);
    }
    template <class T>
    int alloc(T **p, size_t n) {
        if (hipMalloc((void **)p, sizeof(T) * (n ? n : 1)) != hipSuccess) {
            set_error("hipMalloc(%zu) in the batch run failed", sizeof(T) * n);
            return GK_ENOMEM;
        }
        mem.push_back(*p);
        return GK_OK;
    }
};

uint64_t digest(const uint64_t *a, int64_t n, const long long *b) {
    uint64_t h = 1469598103934665603ull;
    for (int64_t i = 0; i < n; ++i) {
        h = (h ^ a[i]) * 1099511628211ull;
        h = (h ^ (uint64_t)b[i]) * 1099511628211ull;
    }
    return h;
}

}  // namespace
}  // namespace gk

extern "C" int gk_batch_tc_ktruss_dev(gk_ctx *ctx, const gk_dcsr *A, int64_t ngraphs, const int64_t *voff, int k,
                                      int use_graph, int reps, uint64_t *triangles, int64_t *truss_edges,
                                      uint8_t *d_keep, double *ms_per_rep, uint64_t *rep_digest,
                                      const gk_opts *opts) {
    GK_CHECK_ARG(ctx && A && voff && triangles && truss_edges, "null argument");
    GK_CHECK_ARG(ngraphs >= 1 && ngraphs < INT32_MAX, "ngraphs out of range");
    GK_CHECK_ARG(k >= 2, "k must be at least 2");
    GK_CHECK_ARG(reps >= 1, "reps must be positive");
    GK_TRY(require_symmetric(A, "A"));
    GK_CHECK_ARG(voff[0] == 0 && voff[ngraphs] == A->m, "voff must run from 0 to A.m");
    for (int64_t g = 0; g < ngraphs; ++g) GK_CHECK_ARG(voff[g + 1] >= voff[g], "voff must not decrease");
    const gk_opts o = resolve_opts(opts);
    // every edge must stay inside its graph
    {
        DevMem V;
        int64_t *dvoff;
        int32_t *bad;
        GK_TRY(V.alloc(&dvoff, ngraphs + 1));
        GK_TRY(V.alloc(&bad, 1));
        GK_HIP(hipMemcpyAsync(dvoff, voff, 8 * (size_t)(ngraphs + 1), hipMemcpyHostToDevice, ctx->stream));
        k_zero_bytes<<<1, 64, 0, ctx->stream>>>(4, (uint8_t *)bad);
        const int g = (int)std::max<int64_t>(1, std::min<int64_t>((A->m + 255) / 256, ctx->num_cus * 16));
        k_check_blocks<<<g, 256, 0, ctx->stream>>>(A->m, ngraphs, dvoff, A->rst, A->col, bad);
        GK_TRY(launch_check("k_check_blocks"));
        int32_t hb = 0;
        GK_HIP(d2h(ctx, &hb, bad, 4));
        GK_CHECK_ARG(hb == 0, "A is not block-diagonal over voff: an edge joins two graphs");
    }
    // structure (not part of the timed step): L, endpoints, entry edge ids
    EdgeIndex X;
    GK_TRY(edge_index(ctx, A, &X));
    const int64_t nL = X.nL;
    // graph g's edges: rows [voff[g], voff[g+1]) of L; block-diagonal A keeps
    // every edge inside its graph
    std::vector<int32_t> eoff_h(ngraphs + 1), order_h(ngraphs);
    {
        DevMem V;
        int64_t *dvoff;
        int32_t *deoff;
        GK_TRY(V.alloc(&dvoff, ngraphs + 1));
        GK_TRY(V.alloc(&deoff, ngraphs + 1));
        GK_HIP(hipMemcpyAsync(dvoff, voff, 8 * (size_t)(ngraphs + 1), hipMemcpyHostToDevice, ctx->stream));
        const int g = (int)std::max<int64_t>(1, std::min<int64_t>((ngraphs + 256) / 256, ctx->num_cus * 4));
        k_gather_offsets<<<g, 256, 0, ctx->stream>>>(ngraphs + 1, dvoff, X.lrst, deoff);
        GK_TRY(launch_check("k_gather_offsets"));
        GK_HIP(d2h(ctx, eoff_h.data(), deoff, 4 * (size_t)(ngraphs + 1)));
    }
    std::iota(order_h.begin(), order_h.end(), 0);
    if (o.batch_order == 0)
        std::stable_sort(order_h.begin(), order_h.end(), [&](int32_t x, int32_t y) {
            return eoff_h[x + 1] - eoff_h[x] > eoff_h[y + 1] - eoff_h[y];
        });
    DevMem D;
    int32_t *eoff, *order, *sup, *buf0, *buf1, *fy, *fx, *fh, *pre, *queue;
    uint8_t *st;
    long long *alive, *tri;  // one allocation: k-truss edges, then triangles, per graph
    GK_TRY(D.alloc(&eoff, ngraphs + 1));
    GK_TRY(D.alloc(&order, ngraphs));
    GK_TRY(D.alloc(&sup, nL));
    GK_TRY(D.alloc(&st, nL));
    GK_TRY(D.alloc(&buf0, nL));
    GK_TRY(D.alloc(&buf1, nL));
    GK_TRY(D.alloc(&fy, nL));
    GK_TRY(D.alloc(&fx, nL));
    GK_TRY(D.alloc(&fh, nL));
    GK_TRY(D.alloc(&pre, nL + ngraphs));
    GK_TRY(D.alloc(&queue, 1));
    GK_TRY(D.alloc(&alive, 2 * (int64_t)ngraphs));
    tri = alive + ngraphs;
    GK_HIP(hipMemcpyAsync(eoff, eoff_h.data(), 4 * (size_t)(ngraphs + 1), hipMemcpyHostToDevice, ctx->stream));
    GK_HIP(hipMemcpyAsync(order, order_h.data(), 4 * (size_t)ngraphs, hipMemcpyHostToDevice, ctx->stream));
    // plan of the support product C<L> = L*L (also outside the step)
    gk_dcsr Ld = view(ctx, A->m, A->m, nL, X.lrst, X.lcol, nullptr);
    Prob P;
    ClassLaunch CL;
    // the union's rows are mostly short (MCA-wave), so that class gets more
    // workgroups than in a single graph's product: the CUs the window classes
    // release when they finish first go to it (4096 graphs: 80.3 -> 72.2 ms)
    gk_opts ob = o;
    if (ob.blocks_mca <= 0) ob.blocks_mca = kBatchMcaBlocksPerCu;
    GK_TRY(prepare_product(ctx, &Ld, &Ld, &Ld, sup, ob, &P, &CL));
    GK_HIP(hipStreamSynchronize(ctx->stream));
    const int thresh = k - 2;
    const int nb = (int)std::max<int64_t>(
        1, std::min<int64_t>(ngraphs, (int64_t)ctx->num_cus * std::max(1, std::min(o.batch_blocks, 32))));
    const int gz = (int)std::max<int64_t>(1, std::min<int64_t>((nL + 255) / 256, ctx->num_cus * 16));
    auto step = [&]() -> int {
        k_zero_bytes<<<gz, 256, 0, ctx->stream>>>((int64_t)nL, st);
        k_zero_bytes<<<1, 64, 0, ctx->stream>>>(4, (uint8_t *)queue);
        GK_TRY(zero_u32(ctx, sup, nL));
        GK_TRY(launch_classes<MODE_SUPPORT>(ctx, P, CL, nullptr));
        k_peel_graphs<<<nb, kPeelThreads, 0, ctx->stream>>>((int32_t)ngraphs, order, eoff, A->rst, A->col, X.aeid,
                                                            X.eu, X.ev, sup, st, thresh, buf0, buf1, fy, fx, fh,
                                                            pre, queue, alive);
        if (d_keep)
            k_keep_flags<<<gz, 256, 0, ctx->stream>>>(A->nz, X.aeid, st, d_keep);
        return launch_check("batch step");
    };
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    auto release = [&] {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    };
    if (use_graph) {
        GK_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        const int rc = step();
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
        graph = g;
        if (rc != GK_OK) { release(); return rc; }
        if (ec != hipSuccess) {
            release();
            set_error("hipStreamEndCapture failed: %s", hipGetErrorString(ec));
            return GK_EHIP;
        }
        if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
            release();
            set_error("hipGraphInstantiate failed");
            return GK_EHIP;
        }
    }
    std::vector<uint64_t> th(ngraphs);
    std::vector<long long> ah(ngraphs);
    Timer t(ctx);
    double ms = 0;
    for (int r = 0; r < reps; ++r) {
        t.start();
        int rc = GK_OK;
        if (use_graph) {
            if (hipGraphLaunch(exec, ctx->stream) != hipSuccess) {
                set_error("hipGraphLaunch failed");
                rc = GK_EHIP;
            }
        } else {
            rc = step();
        }
        t.stop();
        if (rc != GK_OK) { release(); return rc; }
        ms += t.ms();
        if (rep_digest) {
            if (d2h(ctx, th.data(), tri, 8 * (size_t)ngraphs) != hipSuccess ||
                d2h(ctx, ah.data(), alive, 8 * (size_t)ngraphs) != hipSuccess) {
                release();
                set_error("copy of the batch results failed");
                return GK_EHIP;
            }
            rep_digest[r] = digest(th.data(), ngraphs, ah.data());
        }
    }
    release();
    if (d2h(ctx, triangles, tri, 8 * (size_t)ngraphs) != hipSuccess ||
        d2h(ctx, truss_edges, alive, 8 * (size_t)ngraphs) != hipSuccess) {
        set_error("copy of the batch results failed");
        return GK_EHIP;
    }
    if (ms_per_rep) *ms_per_rep = ms / reps;
    ctx->last = gk_timing{};
    ctx->last.total_ms = ms / reps;
    ctx->last.iterations = reps;
    return GK_OK;
}
