// Complement-mask product C<not M> = A*B (GraphKokkos's mxmm_msa_cmask,
// include/matrix_la.h:733-934): C holds every (i, j) that some product
// term A(i,k)*B(k,j) reaches and M does not hold, with the sum of those
// terms (int32, wrap-around), explicit zeros kept, rows in column order.
//
// Unlike C<M> = A*B the output pattern is not known beforehand, so the
// product runs as expand / sort / reduce over 64-bit keys (row << 32 | col):
//   1. k_cm_flop: product terms per row; scanned, they cut the rows into
//      chunks of at most kChunkTerms terms;
//   2. per chunk, k_cm_expand writes one (key, value) per term -- a term at
//      a mask column gets the all-ones key, which sorts last -- then a radix
//      sort of the pairs and a reduce-by-key sum them per (i, j);
//   3. the chunks' results are concatenated and the row pointers come from
//      a count of each row's keys.
// The reference's two passes (symbolic row counts, then values through a
// dense B.n accumulator per thread) become sorts here: HBM-bound key
// traffic instead of a B.n-sized scan per row.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce_by_key.hpp>
#include <rocprim/device/device_scan.hpp>

#include <vector>

#include "gk_device.h"

using namespace gk;

namespace gk {
namespace {

constexpr uint64_t kMasked = ~0ull;
constexpr int64_t kChunkTerms = int64_t(1) << 27;

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
    if (blockIdx.x == 0 && threadIdx.x == 0) flop[m] = 0;
}

// One wave per row of [r0, r1): the row's terms go to [foff[i] - base, ...),
// A entry by A entry, the lanes striding over B's row.
__global__ void k_cm_expand(int64_t r0, int64_t r1, int64_t base, const int32_t *__restrict__ arst,
                            const int32_t *__restrict__ acol, const int32_t *__restrict__ aval,
                            const int32_t *__restrict__ brst, const int32_t *__restrict__ bcol,
                            const int32_t *__restrict__ bval, const int32_t *__restrict__ mrst,
                            const int32_t *__restrict__ mcol, const int64_t *__restrict__ foff,
                            uint64_t *__restrict__ key, uint32_t *__restrict__ val) {
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t i = r0 + (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < r1; i += waves) {
        int64_t out = foff[i] - base;
        const int32_t m0 = mrst[i], m1 = mrst[i + 1];
        for (int32_t t = arst[i]; t < arst[i + 1]; ++t) {
            const int32_t k = acol[t];
            const uint32_t a = aval ? (uint32_t)aval[t] : 1u;
            const int32_t b0 = brst[k], len = brst[k + 1] - b0;
            for (int32_t x = lane; x < len; x += 64) {
                const int32_t j = bcol[b0 + x];
                int32_t l = m0, h = m1;  // j among M's row (sorted)
                while (l < h) {
                    const int32_t mid = l + ((h - l) >> 1);
                    if (mcol[mid] < j) l = mid + 1; else h = mid;
                }
                const bool masked = l < m1 && mcol[l] == j;
                key[out + x] = masked ? kMasked : ((uint64_t)i << 32 | (uint32_t)j);
                val[out + x] = a * (bval ? (uint32_t)bval[b0 + x] : 1u);
            }
            out += len;
        }
    }
}

__global__ void k_cm_split(int64_t n, const uint64_t *__restrict__ key, int32_t *__restrict__ col,
                           int32_t *__restrict__ rowcnt) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        col[p] = (int32_t)(key[p] & 0xffffffffull);
        atomicAdd(&rowcnt[key[p] >> 32], 1);
    }
}

struct Mem {
    std::vector<void *> p;
    ~Mem() {
        for (void *x : p) (void)hipFree(x);
    }
    template <class T>
    int alloc(T **out, int64_t n) {
        if (hipMalloc((void **)out, sizeof(T) * (size_t)(n > 0 ? n : 1)) != hipSuccess) {
            set_error("hipMalloc of %lld bytes for C<not M> failed", (long long)(sizeof(T) * n));
            return GK_ENOMEM;
        }
        p.push_back(*out);
        return GK_OK;
    }
};

int grid_of(gk_ctx *ctx, int64_t n, int64_t per_block) {
    return (int)std::max<int64_t>(1, std::min<int64_t>((n + per_block - 1) / per_block, (int64_t)ctx->num_cus * 16));
}

}  // namespace
}  // namespace gk

extern "C" int gk_masked_spgemm_cmask_dev(gk_ctx *ctx, const gk_dcsr *A, const gk_dcsr *B, const gk_dcsr *M,
                                          gk_dcsr **C) {
    GK_CHECK_ARG(ctx && A && B && M && C, "null argument");
    if (A->n != B->m || A->m != M->m || B->n != M->n) {
        set_error("shape mismatch: A %lldx%lld, B %lldx%lld, M %lldx%lld", (long long)A->m, (long long)A->n,
                  (long long)B->m, (long long)B->n, (long long)M->m, (long long)M->n);
        return GK_EINVAL;
    }
    GK_TRY(require_sorted(M, "M"));
    const int64_t m = A->m;
    Mem T;
    int64_t *flop, *foff;
    GK_TRY(T.alloc(&flop, m + 1));
    GK_TRY(T.alloc(&foff, m + 1));
    k_cm_flop<<<grid_of(ctx, m, 4), 256, 0, ctx->stream>>>(m, A->rst, A->col, B->rst, flop);
    GK_TRY(launch_check("k_cm_flop"));
    {
        size_t tmp = 0;
        GK_HIP(rocprim::exclusive_scan(nullptr, tmp, flop, foff, (int64_t)0, (size_t)(m + 1),
                                       rocprim::plus<int64_t>(), ctx->stream));
        void *t;
        GK_TRY(T.alloc((char **)&t, (int64_t)tmp));
        GK_HIP(rocprim::exclusive_scan(t, tmp, flop, foff, (int64_t)0, (size_t)(m + 1), rocprim::plus<int64_t>(),
                                       ctx->stream));
    }
    std::vector<int64_t> hoff(m + 1);
    GK_HIP(d2h(ctx, hoff.data(), foff, sizeof(int64_t) * (size_t)(m + 1)));
    // row chunks of at most kChunkTerms terms (a longer row is a chunk alone)
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
    uint64_t *key, *skey, *ukey;
    uint32_t *val, *sval, *uval;
    int64_t *nu;
    GK_TRY(T.alloc(&key, maxterms));
    GK_TRY(T.alloc(&skey, maxterms));
    GK_TRY(T.alloc(&ukey, maxterms));
    GK_TRY(T.alloc(&val, maxterms));
    GK_TRY(T.alloc(&sval, maxterms));
    GK_TRY(T.alloc(&uval, maxterms));
    GK_TRY(T.alloc(&nu, 1));
    size_t sort_tmp = 0, red_tmp = 0;
    GK_HIP(rocprim::radix_sort_pairs(nullptr, sort_tmp, key, skey, val, sval, (size_t)maxterms, 0, 64, ctx->stream));
    GK_HIP(rocprim::reduce_by_key(nullptr, red_tmp, skey, sval, (size_t)maxterms, ukey, uval, nu,
                                  rocprim::plus<uint32_t>(), rocprim::equal_to<uint64_t>(), ctx->stream));
    void *stmp, *rtmp;
    GK_TRY(T.alloc((char **)&stmp, (int64_t)sort_tmp));
    GK_TRY(T.alloc((char **)&rtmp, (int64_t)red_tmp));
    // per chunk: unique (key, sum) pairs, kept until the total is known
    std::vector<uint64_t *> ck;
    std::vector<uint32_t *> cv;
    std::vector<int64_t> cn;
    for (size_t c = 1; c < cuts.size(); ++c) {
        const int64_t r0 = cuts[c - 1], r1 = cuts[c], terms = hoff[r1] - hoff[r0];
        if (terms == 0) continue;
        k_cm_expand<<<grid_of(ctx, r1 - r0, 4), 256, 0, ctx->stream>>>(
            r0, r1, hoff[r0], A->rst, A->col, A->val, B->rst, B->col, B->val, M->rst, M->col, foff, key, val);
        GK_TRY(launch_check("k_cm_expand"));
        GK_HIP(rocprim::radix_sort_pairs(stmp, sort_tmp, key, skey, val, sval, (size_t)terms, 0, 64, ctx->stream));
        GK_HIP(rocprim::reduce_by_key(rtmp, red_tmp, skey, sval, (size_t)terms, ukey, uval, nu,
                                      rocprim::plus<uint32_t>(), rocprim::equal_to<uint64_t>(), ctx->stream));
        int64_t hn = 0;
        GK_HIP(d2h(ctx, &hn, nu, sizeof hn));
        uint64_t last = 0;
        GK_HIP(d2h(ctx, &last, ukey + hn - 1, sizeof last));
        if (last == kMasked) --hn;  // the masked terms' key
        if (hn == 0) continue;
        uint64_t *k2;
        uint32_t *v2;
        GK_TRY(T.alloc(&k2, hn));
        GK_TRY(T.alloc(&v2, hn));
        GK_HIP(hipMemcpyAsync(k2, ukey, 8 * (size_t)hn, hipMemcpyDeviceToDevice, ctx->stream));
        GK_HIP(hipMemcpyAsync(v2, uval, 4 * (size_t)hn, hipMemcpyDeviceToDevice, ctx->stream));
        ck.push_back(k2);
        cv.push_back(v2);
        cn.push_back(hn);
    }
    int64_t nz = 0;
    for (int64_t x : cn) nz += x;
    if (nz > INT32_MAX) {
        set_error("C<not M> has %lld entries (int32 CSR limit)", (long long)nz);
        return GK_ERANGE;
    }
    // the result (owned by the returned handle)
    auto *d = new gk_dcsr();
    d->ctx = ctx;
    d->m = M->m;
    d->n = M->n;
    d->nz = nz;
    d->owns = true;
    d->sorted = true;
    auto fail = [&](int rc) {
        free_csr(d);
        delete d;
        return rc;
    };
    if (hipMalloc(&d->rst, 4 * (size_t)(m + 1)) != hipSuccess || hipMalloc(&d->col, 4 * (size_t)std::max<int64_t>(nz, 1)) != hipSuccess ||
        hipMalloc(&d->val, 4 * (size_t)std::max<int64_t>(nz, 1)) != hipSuccess) {
        set_error("hipMalloc of C<not M> failed");
        return fail(GK_ENOMEM);
    }
    int32_t *rowcnt;
    int rc = T.alloc(&rowcnt, m + 1);
    if (rc != GK_OK) return fail(rc);
    if (hipMemsetAsync(rowcnt, 0, 4 * (size_t)(m + 1), ctx->stream) != hipSuccess) return fail(GK_EHIP);
    int64_t at = 0;
    for (size_t c = 0; c < ck.size(); ++c) {
        k_cm_split<<<grid_of(ctx, cn[c], 256), 256, 0, ctx->stream>>>(cn[c], ck[c], d->col + at, rowcnt);
        if (hipMemcpyAsync(d->val + at, cv[c], 4 * (size_t)cn[c], hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess)
            return fail(GK_EHIP);
        at += cn[c];
    }
    if ((rc = launch_check("k_cm_split")) != GK_OK) return fail(rc);
    {
        size_t tmp = 0;
        if (rocprim::exclusive_scan(nullptr, tmp, rowcnt, d->rst, 0, (size_t)(m + 1), rocprim::plus<int32_t>(),
                                    ctx->stream) != hipSuccess)
            return fail(GK_EHIP);
        void *t;
        if ((rc = T.alloc((char **)&t, (int64_t)tmp)) != GK_OK) return fail(rc);
        if (rocprim::exclusive_scan(t, tmp, rowcnt, d->rst, 0, (size_t)(m + 1), rocprim::plus<int32_t>(),
                                    ctx->stream) != hipSuccess)
            return fail(GK_EHIP);
    }
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
        set_error("C<not M> failed on the device");
        return fail(GK_EHIP);
    }
    *C = d;
    return GK_OK;
}

extern "C" int gk_dcsr_download_val(gk_ctx *ctx, const gk_dcsr *d, int32_t *val) {
    GK_CHECK_ARG(ctx && d && (val || d->nz == 0), "null argument");
    GK_CHECK_ARG(d->val || d->nz == 0, "the CSR holds a pattern (no values)");
    if (d->nz) GK_HIP(d2h(ctx, val, d->val, sizeof(int32_t) * (size_t)d->nz));
    return GK_OK;
}
