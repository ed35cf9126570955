// Batched Brandes betweenness centrality on gfx950, sigma and delta in FP64.
//
// GraphKokkos's betweenness_centrality_batch (include/betweenness_centrality.h:406-506)
// keeps a dense n x batch matrix of path counts (Numsp), the frontier of
// every level as a sparse n x batch matrix (Sigmas[d]), advances the
// frontier with a complement-masked product Front' = (A^T * Front) .* !Numsp
// (bc_forward_step, :313-401) and accumulates dependencies backwards with
// W = Nspinv .* Bcu on level d+1 and Bcu += (A * W) .* Numsp on level d
// (bc_backward_step, :182-242); bc(v) = sum_j Bcu(v,j) - batch.
//
// Layout here, for a batch of up to 256 sources (one bit / one column each):
//   S[v][256]  FP64   sigma(v,s) while going forward; overwritten by
//                     W(v,s) = (1/sigma) * (1 + delta) once v's level for s
//                     has been processed backwards
//   vis[v][4]  u64    sources that have reached v
//   F_d[v][4]  u64    sources for which v is at BFS level d (one per level)
//   bc[v]      FP64   sum over sources of delta(v,s)
// One wave per vertex; lane l owns sources 4l..4l+3, so a vertex's 256
// sigmas (2 KiB) are read by one coalesced wave access, and a neighbour's
// frontier bits cost one 8-byte word per 16 lanes.
// Forward (pull over A^T): sigma(i,s) = sum over in-neighbours k with
//   k in F_d for s, for s not yet in vis[i].
// Backward (pull over A): for s with i in F_d,
//   delta(i,s) = sigma(i,s) * sum over out-neighbours w in F_{d+1} for s of W(w,s).
#include <cstring>
#include <memory>

#include "bc.h"
#include "gk_device.h"

using namespace gk;

namespace gk {

int exclusive_scan_i32(gk_ctx *ctx, const int32_t *in, int32_t *out, int64_t n);

namespace {

constexpr int kSrc = bc::kSrc;
// vertices with this many neighbours get a workgroup of kHeavyWaves waves
constexpr int kHeavy = 1024;
constexpr int kHeavyWaves = 16;

__device__ __forceinline__ uint32_t nib_of(const uint64_t *row, int lane) {
    return (uint32_t)(row[lane >> 4] >> ((lane & 15) * 4)) & 0xFu;
}

// Combine the lanes' nibbles of a wave into the 4 words of a row.
__device__ __forceinline__ uint64_t gather_word(uint32_t nib, int lane) {
    uint64_t w = (uint64_t)nib << ((lane & 15) * 4);
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) w |= __shfl_xor(w, o, 16);
    return w;
}

__global__ __launch_bounds__(256) void k_bc_init(int64_t n, uint64_t *__restrict__ vis, uint64_t *__restrict__ f0,
                                                 double *__restrict__ bc, int zero_bc) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < 4 * n; i += (int64_t)gridDim.x * blockDim.x) {
        vis[i] = 0;
        f0[i] = 0;
        if (zero_bc && i < n) bc[i] = 0;
    }
}

__global__ void k_bc_sources(const int32_t *__restrict__ src, int nsrc, uint64_t *__restrict__ vis,
                             uint64_t *__restrict__ f0, double *__restrict__ S) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nsrc) {
        int64_t v = src[s];
        atomicOr((unsigned long long *)&vis[4 * v + (s >> 6)], 1ull << (s & 63));
        atomicOr((unsigned long long *)&f0[4 * v + (s >> 6)], 1ull << (s & 63));
        S[v * kSrc + s] = 1.0;
    }
}

// Sum of S[w][4 lane + b] over the neighbours w = col[t], t in [t0, t1) with
// stride `step`, whose frontier bits (fbits) intersect `mask` at b.  U
// neighbours per step: their ids, then their frontier words, then the
// matching S values are loaded together, so a long list costs about
// deg / (step U) memory round trips.  Neighbours are visited in increasing
// t, so for one (vertex, source) the additions run in list order.
template <int U>
__device__ __forceinline__ void pull_sum(int32_t t0, int32_t t1, int32_t step, const int32_t *__restrict__ col,
                                         const uint64_t *__restrict__ fbits, uint32_t mask,
                                         const double *__restrict__ S, int lane, double *a) {
    if (!mask) return;
    const int word = lane >> 4, sh = (lane & 15) * 4;
    for (int32_t t = t0; t < t1; t += step * U) {
        // Every load of a batch is issued before any is used (the column
        // loads at clamped positions, then the level bits, then the four
        // doubles of each neighbour with a matching source, predicated with
        // no branch on loaded data between them), so U neighbours are in
        // flight per lane.  The sums add the set sources in neighbour order,
        // as a loop over the neighbours one by one would.
        int32_t w[U];
        uint32_t f[U];
#pragma unroll
        for (int u = 0; u < U; ++u) w[u] = col[min(t + u * step, t1 - 1)];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t b = (uint32_t)(fbits[4 * (int64_t)w[u] + word] >> sh) & mask;
            f[u] = t + u * step < t1 ? b : 0u;
        }
        double2 lo[U], hi[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            lo[u] = make_double2(0.0, 0.0);
            hi[u] = lo[u];
            if (f[u]) {
                const double2 *sw = reinterpret_cast<const double2 *>(S + (int64_t)w[u] * kSrc + 4 * lane);
                lo[u] = sw[0];
                hi[u] = sw[1];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (f[u] & 1) a[0] += lo[u].x;
            if (f[u] & 2) a[1] += lo[u].y;
            if (f[u] & 4) a[2] += hi[u].x;
            if (f[u] & 8) a[3] += hi[u].y;
        }
    }
}

// Forward finish for vertex i: sigma of newly reached sources, frontier and
// visited bits.  Returns the number of (vertex, source) pairs reached.
__device__ __forceinline__ uint32_t forward_finish(int64_t i, uint32_t unv, const double *a, uint64_t *fn,
                                                   uint64_t *vis, double *S, int lane) {
    uint32_t got = 0;
    double *si = S + i * kSrc + 4 * lane;
#pragma unroll
    for (int b = 0; b < 4; ++b)
        if ((unv & (1u << b)) && a[b] != 0) {
            si[b] = a[b];
            got |= 1u << b;
        }
    const uint64_t w = gather_word(got, lane);
    if ((lane & 15) == 0) {
        fn[4 * i + (lane >> 4)] = w;
        if (w) vis[4 * i + (lane >> 4)] |= w;
    }
    return __popc(got);
}

// Backward finish for vertex i at level d: delta = sigma * sum, W = (1/sigma)
// * (1 + delta) in place of sigma; returns the wave's sum of delta.
__device__ __forceinline__ double backward_finish(int64_t i, uint32_t nd, const double *a, double *S, int lane) {
    double *si = S + i * kSrc + 4 * lane;
    double dsum = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        if (nd & (1u << b)) {
            const double sigma = si[b];
            const double delta = a[b] * sigma;
            si[b] = (1.0 / sigma) * (1.0 + delta);
            dsum += delta;
        }
    }
    return dev::wave_sum(dsum);
}

// Level d -> d+1 for vertices with fewer than kHeavy in-neighbours, one wave
// per vertex (heavier ones: k_bc_heavy_chunk).  cnt accumulates the number
// of (vertex, source) pairs reached.
// Eight waves per SIMD (64 VGPRs, one under the 65 it would take): a level
// is a chain of dependent loads per vertex.  Together with the heavy-chunk
// kernel's eight waves: R-MAT scale 20, 4 x 256 sources, 106.5 -> 99.8 ms.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_bc_forward(int64_t n, const int32_t *__restrict__ trst,
                                                    const int32_t *__restrict__ tcol,
                                                    const uint64_t *__restrict__ fd, uint64_t *__restrict__ fn,
                                                    uint64_t *__restrict__ vis, double *__restrict__ S,
                                                    unsigned long long *__restrict__ cnt) {
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    uint32_t reached = 0;
    for (int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < n; i += waves) {
        const int32_t t0 = trst[i], t1 = trst[i + 1];
        if (t1 - t0 >= kHeavy) continue;
        const uint32_t unv = ~nib_of(vis + 4 * i, lane) & 0xFu;
        double a[4] = {0, 0, 0, 0};
        if (__ballot(unv != 0)) pull_sum<4>(t0, t1, 1, tcol, fd, unv, S, lane, a);
        reached += forward_finish(i, unv, a, fn, vis, S, lane);
    }
    unsigned long long r = dev::wave_sum((unsigned long long)reached);
    if (lane == 0 && r) atomicAdd(cnt, r);
}

// Heavy vertices in chunks: a vertex with kHeavy or more neighbours is cut
// into chunks of kChunk neighbours, one workgroup of kHeavyWaves waves per
// chunk, so a hub's sum is spread over many workgroups instead of being the
// long serial tail of one.  chunks[c] = (vertex, first, end, several): a
// vertex with one chunk is finished by that chunk's workgroup; the chunks
// of a vertex with several write their 256 partial sums to part[c], and
// k_bc_heavy_finish adds them in chunk order (deterministic sums).
// multi[x] = (vertex, first chunk, chunk count) for those vertices.
constexpr int kChunk = 2048;

// One block: the chunks of every heavy vertex (heavy[0, nh)).
__global__ __launch_bounds__(1024) void k_heavy_chunks(int32_t nh, const int32_t *__restrict__ heavy,
                                                       const int32_t *__restrict__ rst, int4 *__restrict__ chunks,
                                                       int4 *__restrict__ multi, int32_t *__restrict__ counts) {
    __shared__ int32_t ws[2][1024 / 64];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int32_t per = (nh + 1023) / 1024;
    const int32_t h0 = min(nh, tid * per), h1 = min(nh, h0 + per);
    int32_t nc = 0, nm = 0;
    for (int32_t h = h0; h < h1; ++h) {
        const int32_t d = rst[heavy[h] + 1] - rst[heavy[h]];
        const int32_t k = (d + kChunk - 1) / kChunk;
        nc += k;
        nm += k > 1;
    }
    const int32_t ic = dev::wave_incl_scan(nc), im = dev::wave_incl_scan(nm);
    if (lane == 63) {
        ws[0][w] = ic;
        ws[1][w] = im;
    }
    __syncthreads();
    int32_t oc = ic - nc, om = im - nm;
    for (int x = 0; x < w; ++x) {
        oc += ws[0][x];
        om += ws[1][x];
    }
    for (int32_t h = h0; h < h1; ++h) {
        const int32_t v = heavy[h], b = rst[v], e = rst[v + 1];
        const int32_t k = (e - b + kChunk - 1) / kChunk;
        if (k > 1) multi[om++] = make_int4(v, oc, k, 0);
        for (int32_t j = 0; j < k; ++j) chunks[oc++] = make_int4(v, b + j * kChunk, min(e, b + (j + 1) * kChunk), k > 1);
    }
    if (tid == 1023) {
        counts[0] = oc;
        counts[1] = om;
    }
}

// The sources of vertex i that a sweep works on: forward, those that have
// not reached i; backward, those for which i is on level d.
template <bool FWD>
__device__ __forceinline__ uint32_t sweep_mask(int64_t i, const uint64_t *__restrict__ vis,
                                               const uint64_t *__restrict__ fd, int lane) {
    return FWD ? ~nib_of(vis + 4 * i, lane) & 0xFu : nib_of(fd + 4 * i, lane);
}

// Forward (FWD): sums sigma over in-neighbours on level d (fd) into the new
// level fn; backward: sums W over out-neighbours on level d+1 (fn).  Eight
// waves per SIMD, so two 16-wave workgroups share a CU (at the cost of a few
// registers spilled: 12 B per lane backward, 36 B forward).
template <bool FWD>
__global__ __launch_bounds__(64 * kHeavyWaves) __attribute__((amdgpu_waves_per_eu(8))) void k_bc_heavy_chunk(
    int32_t nc, const int4 *__restrict__ chunks, const int32_t *__restrict__ col, const uint64_t *__restrict__ fd,
    uint64_t *__restrict__ fn, uint64_t *__restrict__ vis, double *__restrict__ S, double *__restrict__ bc,
    unsigned long long *__restrict__ cnt, double *__restrict__ part) {
    __shared__ double psum[kHeavyWaves][kSrc];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int32_t c = blockIdx.x; c < nc; c += gridDim.x) {
        const int4 ch = chunks[c];
        const int64_t i = ch.x;
        const uint32_t mask = sweep_mask<FWD>(i, vis, fd, lane);
        if (!__syncthreads_or(mask != 0)) {
            // every source has reached i: i is on no new level, and its
            // frontier word is written all the same (the level buffers are
            // reused from batch to batch); a vertex with several chunks gets
            // it from k_bc_heavy_finish
            if (FWD && !ch.w && threadIdx.x < 4) fn[4 * i + threadIdx.x] = 0;
            continue;
        }
        double a[4] = {0, 0, 0, 0};
        pull_sum<4>(ch.y + wv, ch.z, kHeavyWaves, col, FWD ? fd : fn, mask, S, lane, a);
#pragma unroll
        for (int b = 0; b < 4; ++b) psum[wv][4 * lane + b] = a[b];
        __syncthreads();
        if (wv == 0) {
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                double s = 0;
                for (int x = 0; x < kHeavyWaves; ++x) s += psum[x][4 * lane + b];
                a[b] = s;
            }
            if (ch.w) {
#pragma unroll
                for (int b = 0; b < 4; ++b) part[(int64_t)c * kSrc + 4 * lane + b] = a[b];
            } else if (FWD) {
                const uint32_t got = forward_finish(i, mask, a, fn, vis, S, lane);
                const unsigned long long r = dev::wave_sum((unsigned long long)got);
                if (lane == 0 && r) atomicAdd(cnt, r);
            } else {
                const double dsum = backward_finish(i, mask, a, S, lane);
                if (lane == 0) bc[i] += dsum;
            }
        }
        __syncthreads();
    }
}

// The vertices with several chunks: their chunks' partial sums in chunk
// order, then the level's finish; one wave per vertex.
template <bool FWD>
__global__ __launch_bounds__(256) void k_bc_heavy_finish(int32_t nm, const int4 *__restrict__ multi,
                                                         const double *__restrict__ part,
                                                         const uint64_t *__restrict__ fd, uint64_t *__restrict__ fn,
                                                         uint64_t *__restrict__ vis, double *__restrict__ S,
                                                         double *__restrict__ bc, unsigned long long *__restrict__ cnt) {
    const int lane = threadIdx.x & 63;
    const int32_t waves = gridDim.x * (blockDim.x >> 6);
    for (int32_t x = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); x < nm; x += waves) {
        const int4 mv = multi[x];
        const int64_t i = mv.x;
        const uint32_t mask = sweep_mask<FWD>(i, vis, fd, lane);
        if (!__ballot(mask != 0)) {
            if (FWD && lane < 4) fn[4 * i + lane] = 0;  // on no new level
            continue;
        }
        double a[4] = {0, 0, 0, 0};
        for (int32_t j = 0; j < mv.z; ++j) {
            const double *p = part + (int64_t)(mv.y + j) * kSrc + 4 * lane;
#pragma unroll
            for (int b = 0; b < 4; ++b) a[b] += p[b];
        }
        if (FWD) {
            const uint32_t got = forward_finish(i, mask, a, fn, vis, S, lane);
            const unsigned long long r = dev::wave_sum((unsigned long long)got);
            if (lane == 0 && r) atomicAdd(cnt, r);
        } else {
            const double dsum = backward_finish(i, mask, a, S, lane);
            if (lane == 0) bc[i] += dsum;
        }
    }
}

// Deepest level: delta = 0, so W = 1/sigma.
__global__ __launch_bounds__(256) void k_bc_last(int64_t n, const uint64_t *__restrict__ fd, double *__restrict__ S) {
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < n; i += waves) {
        const uint32_t nd = nib_of(fd + 4 * i, lane);
        if (!nd) continue;
        double *si = S + i * kSrc + 4 * lane;
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if (nd & (1u << b)) si[b] = (1.0 / si[b]) * 1.0;
    }
}

// Level d (sources in F_d) from level d+1 (F_{d+1}, already holding W), for
// vertices with fewer than kHeavy out-neighbours.
__global__ __launch_bounds__(256) void k_bc_backward(int64_t n, const int32_t *__restrict__ rst,
                                                     const int32_t *__restrict__ col,
                                                     const uint64_t *__restrict__ fd,
                                                     const uint64_t *__restrict__ fn, double *__restrict__ S,
                                                     double *__restrict__ bc) {
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < n; i += waves) {
        const uint32_t nd = nib_of(fd + 4 * i, lane);
        if (!__ballot(nd != 0)) continue;
        const int32_t t0 = rst[i], t1 = rst[i + 1];
        if (t1 - t0 >= kHeavy) continue;
        double a[4] = {0, 0, 0, 0};
        pull_sum<4>(t0, t1, 1, col, fn, nd, S, lane, a);
        const double dsum = backward_finish(i, nd, a, S, lane);
        if (lane == 0) bc[i] += dsum;
    }
}

// Vertices with at least kHeavy entries in `rst` rows.
// A source row that lists a neighbour v c > 1 times (a Matrix Market file
// may repeat an edge) puts c copies of (v, s) into the reference's first
// frontier (transpose(A.extractRows), include/betweenness_centrality.h:447),
// and bc_backward_step (:236-239) adds (A*W)*Numsp once per copy, so v's
// dependency on s counts c times.  sigma(v,s) = c there, so
// delta = c * W(v,s) - 1 and the extra copies add (c - 1) * delta.
// t holds (source column, v, c) triples.
__global__ void k_bc_dup_fix(int32_t np, const int32_t *__restrict__ t, const double *__restrict__ S,
                             double *__restrict__ bc) {
    const int32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < np) {
        const int32_t j = t[3 * p], v = t[3 * p + 1], c = t[3 * p + 2];
        const double delta = (double)c * S[(int64_t)v * kSrc + j] - 1.0;
        atomicAdd(&bc[v], (double)(c - 1) * delta);
    }
}

__global__ void k_heavy_list(int64_t n, const int32_t *__restrict__ rst, int32_t *__restrict__ out,
                             int32_t *__restrict__ nout) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (rst[i + 1] - rst[i] >= kHeavy) out[atomicAdd(nout, 1)] = (int32_t)i;
}

// A^T of a pattern CSR (counting sort, as transpose in include/matrix_la.h:155-179).
__global__ void k_tcount(int64_t nz, const int32_t *__restrict__ col, int32_t *__restrict__ cnt) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < nz; p += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&cnt[col[p]], 1);
}
__global__ void k_tfill(int64_t m, const int32_t *__restrict__ rst, const int32_t *__restrict__ col,
                        int32_t *__restrict__ cur, int32_t *__restrict__ tcol) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
        for (int32_t p = rst[i]; p < rst[i + 1]; ++p) tcol[atomicAdd(&cur[col[p]], 1)] = (int32_t)i;
}
int grid_waves(gk_ctx *ctx, int64_t rows) {
    int64_t g = (rows + 3) / 4;
    int64_t cap = (int64_t)ctx->num_cus * 16;
    return (int)std::max<int64_t>(1, std::min(g, cap));
}

// Repeated neighbours in the rows of the sources, per batch of kSrc
// sources: (column in the batch, neighbour, multiplicity) for every
// multiplicity above 1.  Only patterns whose rows are not strictly
// increasing can repeat a column.
int repeated_neighbours(gk_ctx *ctx, const gk_dcsr *A, const int32_t *sources, int64_t nsrc,
                        std::vector<std::vector<int32_t>> *per_batch) {
    per_batch->assign((size_t)((nsrc + kSrc - 1) / kSrc), {});
    if (A->sorted || nsrc == 0) return GK_OK;
    std::vector<int32_t> rst(A->m + 1), row;
    GK_HIP(d2h(ctx, rst.data(), A->rst, 4 * (size_t)(A->m + 1)));
    for (int64_t s = 0; s < nsrc; ++s) {
        const int32_t v0 = sources[s], b = rst[v0], e = rst[v0 + 1];
        if (e - b < 2) continue;
        row.resize(e - b);
        GK_HIP(d2h(ctx, row.data(), A->col + b, 4 * (size_t)(e - b)));
        std::sort(row.begin(), row.end());
        for (size_t q = 0; q < row.size();) {
            size_t r = q;
            while (r < row.size() && row[r] == row[q]) ++r;
            if (r - q > 1 && row[q] != v0) {
                auto &t = (*per_batch)[s / kSrc];
                t.push_back((int32_t)(s % kSrc));
                t.push_back(row[q]);
                t.push_back((int32_t)(r - q));
            }
            q = r;
        }
    }
    return GK_OK;
}

// Buffers of one gk_bc_batch_dev call: each one hipMalloc'd, all freed when
// the call ends.
struct DevPool final : bc::Pool {
    int64_t n;
    std::vector<void *> allocs;
    std::vector<uint64_t *> F;
    explicit DevPool(int64_t n_) : n(n_) {}
    ~DevPool() override { for (void *p : allocs) (void)hipFree(p); }
    int get(size_t bytes, void **p) override {
        if (hipMalloc(p, bytes) != hipSuccess) {
            set_error("hipMalloc(%zu) in BC failed", bytes);
            return GK_ENOMEM;
        }
        allocs.push_back(*p);
        return GK_OK;
    }
    int level(int d, uint64_t **p) override {
        while ((int)F.size() <= d) {
            uint64_t *f;
            GK_TRY(alloc(&f, (size_t)n * 4));
            F.push_back(f);
        }
        *p = F[d];
        return GK_OK;
    }
};

}  // namespace

namespace bc {

// The heavy-vertex kernels run on an auxiliary stream beside the light ones:
// the two write disjoint vertices (their sigma or W, frontier and visited
// bits, bc) and read only vertices of the neighbouring level.
int sweeps(gk_ctx *ctx, const gk_dcsr *A, const int32_t *trst, const int32_t *tcol, const int32_t *sources,
           int64_t nsrc, double *d_bc, Pool &P, const std::vector<std::vector<int32_t>> *dups, bool timed) {
    const int64_t n = A->m;
    double *S;
    uint64_t *vis;
    int32_t *dsrc;
    unsigned long long *dcnt;
    int32_t *hin, *hout, *nh;
    GK_TRY(P.alloc(&S, (size_t)n * kSrc));
    GK_TRY(P.alloc(&vis, (size_t)n * 4));
    GK_TRY(P.alloc(&dsrc, kSrc));
    GK_TRY(P.alloc(&dcnt, 1));
    GK_TRY(P.alloc(&hin, n));
    GK_TRY(P.alloc(&hout, n));
    GK_TRY(P.alloc(&nh, 2));
    int32_t *ddup = nullptr;
    size_t dup_max = 0;
    if (dups)
        for (auto &t : *dups) dup_max = std::max(dup_max, t.size());
    if (dup_max) GK_TRY(P.alloc(&ddup, dup_max));
    GK_HIP(hipMemsetAsync(nh, 0, 8, ctx->stream));
    {
        const int g = (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, ctx->num_cus * 16));
        k_heavy_list<<<g, 256, 0, ctx->stream>>>(n, trst, hin, nh);
        k_heavy_list<<<g, 256, 0, ctx->stream>>>(n, A->rst, hout, nh + 1);
        GK_TRY(launch_check("k_heavy_list"));
    }
    unsigned long long *h;
    GK_TRY(pinned(ctx, 32, (void **)&h));
    GK_HIP(hipMemcpyAsync(h, nh, 8, hipMemcpyDeviceToHost, ctx->stream));
    GK_HIP(hipStreamSynchronize(ctx->stream));
    int32_t nheavy[2];
    std::memcpy(nheavy, h, 8);
    // the heavy vertices' chunks: [0] in-neighbours (forward), [1]
    // out-neighbours (backward); a vertex's chunks are at most its
    // neighbours / kChunk + 1, and the heavy vertices' neighbours at most
    // A's entries
    int4 *chunks[2] = {nullptr, nullptr}, *multi[2] = {nullptr, nullptr};
    double *part[2] = {nullptr, nullptr};
    int32_t nchunk[2] = {0, 0}, nmulti[2] = {0, 0}, *dcounts;
    GK_TRY(P.alloc(&dcounts, 4));
    for (int d = 0; d < 2; ++d) {
        if (!nheavy[d]) continue;
        const size_t cap = (size_t)(A->nz / kChunk) + (size_t)nheavy[d] + 1;
        GK_TRY(P.alloc(&chunks[d], cap));
        GK_TRY(P.alloc(&multi[d], nheavy[d]));
        GK_TRY(P.alloc(&part[d], cap * kSrc));
        k_heavy_chunks<<<1, 1024, 0, ctx->stream>>>(nheavy[d], d ? hout : hin, d ? A->rst : trst, chunks[d],
                                                     multi[d], dcounts + 2 * d);
        GK_TRY(launch_check("k_heavy_chunks"));
    }
    if (nheavy[0] || nheavy[1]) {
        GK_HIP(hipMemcpyAsync(h, dcounts, 16, hipMemcpyDeviceToHost, ctx->stream));
        GK_HIP(hipStreamSynchronize(ctx->stream));
        int32_t c4[4];
        std::memcpy(c4, h, 16);
        for (int d = 0; d < 2; ++d)
            if (nheavy[d]) {
                nchunk[d] = c4[2 * d];
                nmulti[d] = c4[2 * d + 1];
            }
    }
    uint64_t *f0;
    GK_TRY(P.level(0, &f0));
    const int gw = grid_waves(ctx, n);
    const int gi = (int)std::max<int64_t>(1, std::min<int64_t>((4 * n + 255) / 256, ctx->num_cus * 16));
    hipStream_t hs = ctx->aux[0];
    // fork the heavy-vertex stream from the context stream / join it back
    auto fork = [&]() -> int {
        GK_HIP(hipEventRecord(ctx->fork, ctx->stream));
        GK_HIP(hipStreamWaitEvent(hs, ctx->fork, 0));
        return GK_OK;
    };
    auto join = [&]() -> int {
        GK_HIP(hipEventRecord(ctx->join[0], hs));
        GK_HIP(hipStreamWaitEvent(ctx->stream, ctx->join[0], 0));
        return GK_OK;
    };
    int64_t levels_total = 0;
    double fwd_ms = 0, bwd_ms = 0;
    for (int64_t b0 = 0; b0 < nsrc || (b0 == 0 && nsrc == 0); b0 += kSrc) {
        const int ns = (int)std::min<int64_t>(kSrc, nsrc - b0);
        k_bc_init<<<gi, 256, 0, ctx->stream>>>(n, vis, f0, d_bc, b0 == 0);
        GK_TRY(launch_check("k_bc_init"));
        if (ns <= 0) break;
        GK_HIP(hipMemcpyAsync(dsrc, sources + b0, 4 * ns, hipMemcpyHostToDevice, ctx->stream));
        k_bc_sources<<<1, 256, 0, ctx->stream>>>(dsrc, ns, vis, f0, S);
        GK_TRY(launch_check("k_bc_sources"));
        // forward
        std::unique_ptr<Timer> tf(timed ? new Timer(ctx) : nullptr);
        if (tf) tf->start();
        int d = 0;
        for (;;) {
            uint64_t *fd, *fn;
            GK_TRY(P.level(d, &fd));
            GK_TRY(P.level(d + 1, &fn));
            GK_HIP(hipMemsetAsync(dcnt, 0, 8, ctx->stream));
            if (nheavy[0]) {
                GK_TRY(fork());
                k_bc_heavy_chunk<true><<<std::min(nchunk[0], ctx->num_cus * 2), 64 * kHeavyWaves, 0, hs>>>(
                    nchunk[0], chunks[0], tcol, fd, fn, vis, S, d_bc, dcnt, part[0]);
                if (nmulti[0])
                    k_bc_heavy_finish<true><<<std::max(1, std::min((nmulti[0] + 3) / 4, ctx->num_cus * 4)), 256, 0,
                                               hs>>>(nmulti[0], multi[0], part[0], fd, fn, vis, S, d_bc, dcnt);
            }
            k_bc_forward<<<gw, 256, 0, ctx->stream>>>(n, trst, tcol, fd, fn, vis, S, dcnt);
            if (nheavy[0]) GK_TRY(join());
            GK_TRY(launch_check("k_bc_forward"));
            GK_HIP(hipMemcpyAsync(h, dcnt, 8, hipMemcpyDeviceToHost, ctx->stream));
            GK_HIP(hipStreamSynchronize(ctx->stream));
            if (h[0] == 0) break;
            ++d;
        }
        if (tf) {
            tf->stop();
            fwd_ms += tf->ms();
        }
        // d = deepest non-empty level
        std::unique_ptr<Timer> tb(timed ? new Timer(ctx) : nullptr);
        if (tb) tb->start();
        if (d >= 1) {
            uint64_t *fd, *fn;
            GK_TRY(P.level(d, &fd));
            k_bc_last<<<gw, 256, 0, ctx->stream>>>(n, fd, S);
            GK_TRY(launch_check("k_bc_last"));
            for (int l = d - 1; l >= 1; --l) {
                GK_TRY(P.level(l, &fd));
                GK_TRY(P.level(l + 1, &fn));
                if (nheavy[1]) {
                    GK_TRY(fork());
                    k_bc_heavy_chunk<false><<<std::min(nchunk[1], ctx->num_cus * 2), 64 * kHeavyWaves, 0, hs>>>(
                        nchunk[1], chunks[1], A->col, fd, fn, nullptr, S, d_bc, nullptr, part[1]);
                    if (nmulti[1])
                        k_bc_heavy_finish<false><<<std::max(1, std::min((nmulti[1] + 3) / 4, ctx->num_cus * 4)), 256,
                                                    0, hs>>>(nmulti[1], multi[1], part[1], fd, fn, nullptr, S, d_bc,
                                                             nullptr);
                }
                k_bc_backward<<<gw, 256, 0, ctx->stream>>>(n, A->rst, A->col, fd, fn, S, d_bc);
                if (nheavy[1]) GK_TRY(join());
                GK_TRY(launch_check("k_bc_backward"));
            }
            if (dups) {
                const auto &t = (*dups)[b0 / kSrc];
                if (!t.empty()) {
                    const int32_t np = (int32_t)(t.size() / 3);
                    GK_HIP(hipMemcpyAsync(ddup, t.data(), 4 * t.size(), hipMemcpyHostToDevice, ctx->stream));
                    k_bc_dup_fix<<<(np + 255) / 256, 256, 0, ctx->stream>>>(np, ddup, S, d_bc);
                    GK_TRY(launch_check("k_bc_dup_fix"));
                }
            }
        }
        if (tb) {
            tb->stop();
            bwd_ms += tb->ms();
        }
        levels_total += d;
    }
    ctx->last.kernel_ms = fwd_ms + bwd_ms;
    ctx->last.select_ms = fwd_ms;  // forward share (backward = kernel_ms - select_ms)
    ctx->last.iterations = levels_total;
    return GK_OK;
}

int check_args(gk_ctx *ctx, const gk_dcsr *A, const gk_dcsr *AT, const int32_t *sources, int64_t nsrc,
                  const double *d_bc) {
    GK_CHECK_ARG(ctx && A && d_bc && (sources || nsrc == 0), "null argument");
    GK_CHECK_ARG(A->m == A->n, "non-square matrix; BC is only for square matrices");
    GK_CHECK_ARG(!AT || (AT->m == A->n && AT->nz == A->nz), "AT does not match A");
    for (int64_t s = 0; s < nsrc; ++s)
        GK_CHECK_ARG(sources[s] >= 0 && sources[s] < A->m, "source %lld out of range", (long long)sources[s]);
    return GK_OK;
}

}  // namespace bc
}  // namespace gk

extern "C" int gk_bc_batch_dev(gk_ctx *ctx, const gk_dcsr *A, const gk_dcsr *AT, const int32_t *sources, int64_t nsrc,
                               double *d_bc, const gk_opts *opts) {
    GK_TRY(bc::check_args(ctx, A, AT, sources, nsrc, d_bc));
    (void)opts;
    const int64_t n = A->m;
    DevPool D(n);
    const int32_t *trst = AT ? AT->rst : nullptr, *tcol = AT ? AT->col : nullptr;
    // a symmetric pattern is its own transpose (row order inside A^T does
    // not matter to the forward sums)
    if (!AT && A->symmetric) {
        trst = A->rst;
        tcol = A->col;
    }
    Timer total(ctx);
    total.start();
    if (!trst) {
        int32_t *cnt, *tr, *tc, *cur;
        GK_TRY(D.alloc(&cnt, n + 1));
        GK_TRY(D.alloc(&tr, n + 1));
        GK_TRY(D.alloc(&tc, A->nz));
        GK_TRY(D.alloc(&cur, n + 1));
        GK_HIP(hipMemsetAsync(cnt, 0, 4 * (n + 1), ctx->stream));
        int g = (int)std::max<int64_t>(1, std::min<int64_t>((A->nz + 255) / 256, ctx->num_cus * 16));
        if (A->nz) k_tcount<<<g, 256, 0, ctx->stream>>>(A->nz, A->col, cnt);
        GK_TRY(exclusive_scan_i32(ctx, cnt, tr, n + 1));
        GK_HIP(hipMemcpyAsync(cur, tr, 4 * (n + 1), hipMemcpyDeviceToDevice, ctx->stream));
        int gm = (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, ctx->num_cus * 16));
        // Row order of A^T does not matter: forward sums add integers
        // (exact in FP64 below 2^53).
        k_tfill<<<gm, 256, 0, ctx->stream>>>(n, A->rst, A->col, cur, tc);
        GK_TRY(launch_check("transpose"));
        trst = tr;
        tcol = tc;
    }
    std::vector<std::vector<int32_t>> dups;
    GK_TRY(repeated_neighbours(ctx, A, sources, nsrc, &dups));
    ctx->last = gk_timing{};
    GK_TRY(bc::sweeps(ctx, A, trst, tcol, sources, nsrc, d_bc, D, &dups, true));
    total.stop();
    ctx->last.total_ms = total.ms();
    return GK_OK;
}

extern "C" int gk_bc_batch_f64(gk_ctx *ctx, const gk_csr *A, const int32_t *sources, int64_t nsrc, double *bc,
                               const gk_opts *opts) {
    GK_CHECK_ARG(ctx && A && bc, "null argument");
    gk_dcsr d;
    GK_TRY(upload_csr(ctx, A, &d));
    double *db = nullptr;
    if (hipMalloc(&db, 8 * (A->m ? A->m : 1)) != hipSuccess) { free_csr(&d); set_error("hipMalloc failed"); return GK_ENOMEM; }
    int rc = gk_bc_batch_dev(ctx, &d, nullptr, sources, nsrc, db, opts);
    if (rc == GK_OK && A->m && d2h(ctx, bc, db, 8 * A->m) != hipSuccess) {
        set_error("copy of bc failed");
        rc = GK_EHIP;
    }
    (void)hipFree(db);
    free_csr(&d);
    return rc;
}
