// graphkokkos/matrix_la.h -- drop-in for the masked products of GraphKokkos's
// include/matrix_la.h: mxmm_msa and mxmm_mca keep their names and
// signatures (include/matrix_la.h:80-123) and run C<M> = A*B on the GPU
// through gk_masked_spgemm_i32.  mxmm_gpu is the same operator under a name
// of its own; every one of them converts to the reference's product
// pointer type mxmOp<int> (src/main.cpp:13-14).  mxmm_msa_cmask, the
// complement-mask product, runs through gk_masked_spgemm_cmask_dev.
//
// Semantics (include/matrix_la.h:478-491, 562-574): C takes M's shape and
// pattern (Rst and Col copied from M), C.Val[p] = sum_k A(i,k)*B(k,j) for
// M's entry p = (i,j), explicit zeros kept, int32 arithmetic with wrap-around.
// Masks with rows in file order or repeated columns give every copy of a
// column the column's value, as the reference's MSA does.  isParallel is
// accepted for signature compatibility; the GPU product is always the
// parallel one.  Accumulator choice comes from gk::options(): mxmm_msa and
// mxmm_mca leave the per-row choice to the library's selector unless
// gk::options() forces one.
#pragma once

#include <type_traits>
#include <vector>

#include "matrix.h"

template <typename T>
using mxmOp = void (*)(bool, const spMtx<T> &, const spMtx<T> &, const spMtx<T> &, spMtx<T> &);

namespace gk {
namespace detail {

// Sorted, duplicate-free copy of a mask pattern plus the copy's entry for
// every entry of the original; empty `pos` when the mask is already
// canonical and is passed through unchanged.
struct CanonicalMask {
    std::vector<int32_t> rst, col;
    std::vector<int64_t> pos;
    gk_csr csr{};
};

template <typename U>
CanonicalMask canonical_mask(const spMtx<U> &M) {
    CanonicalMask c;
    c.csr = M.csr(false);
    bool canonical = true;
    for (size_t i = 0; i < M.m && canonical; ++i)
        for (int p = M.Rst(i) + 1; p < M.Rst(i + 1); ++p)
            if (M.Col(p) <= M.Col(p - 1)) { canonical = false; break; }
    if (canonical) return c;
    c.rst.assign(M.m + 1, 0);
    c.col.reserve(M.nz);
    c.pos.resize(M.nz);
    std::vector<std::pair<int32_t, int32_t>> row;   // (column, entry)
    for (size_t i = 0; i < M.m; ++i) {
        row.clear();
        for (int p = M.Rst(i); p < M.Rst(i + 1); ++p) row.emplace_back(M.Col(p), p);
        std::sort(row.begin(), row.end());
        for (size_t t = 0; t < row.size(); ++t) {
            if (t == 0 || row[t].first != row[t - 1].first) c.col.push_back(row[t].first);
            c.pos[row[t].second] = (int64_t)c.col.size() - 1;
        }
        c.rst[i + 1] = (int32_t)c.col.size();
    }
    c.csr.nz = (int64_t)c.col.size();
    c.csr.rst = c.rst.data();
    c.csr.col = c.col.data();
    return c;
}

template <typename T, typename U>
void masked_product(const spMtx<T> &A, const spMtx<T> &B, const spMtx<U> &M, spMtx<T> &C) {
    static_assert(std::is_same<T, int>::value, "the GPU masked product computes int32 values (spMtx<int>)");
    if (A.n != B.m) throw error(GK_EINVAL, "masked product: A.n != B.m");
    const gk_csr a = A.csr(), b = B.csr();
    const CanonicalMask cm = canonical_mask(M);
    std::vector<int32_t> vals((size_t)cm.csr.nz);
    check(gk_masked_spgemm_i32(context(), &a, &b, &cm.csr, vals.data(), &options()), "gk_masked_spgemm_i32");
    spMtx<T> R(M.m, M.n, M.nz);
    std::memcpy(R.matcode, C.matcode, sizeof(MM_typecode));
    std::copy(M.Rst.data(), M.Rst.data() + M.m + 1, R.Rst.data());
    std::copy(M.Col.data(), M.Col.data() + M.nz, R.Col.data());
    for (size_t p = 0; p < M.nz; ++p) R.Val(p) = vals[cm.pos.empty() ? p : (size_t)cm.pos[p]];
    C = std::move(R);
}

// C<not M> = A*B: the output pattern comes from the device, so A, B and the
// canonical mask go up as device CSRs and the library-built result comes
// back whole (rows, columns, values).
struct DeviceCsr {
    gk_dcsr *d = nullptr;
    ~DeviceCsr() {
        if (d) gk_dcsr_free(d);
    }
};

template <typename T, typename U>
void complement_product(const spMtx<T> &A, const spMtx<T> &B, const spMtx<U> &M, spMtx<T> &C) {
    static_assert(std::is_same<T, int>::value, "the GPU masked product computes int32 values (spMtx<int>)");
    if (A.n != B.m) throw error(GK_EINVAL, "complement-mask product: A.n != B.m");
    const gk_csr a = A.csr(), b = B.csr();
    const CanonicalMask cm = canonical_mask(M);
    DeviceCsr da, db, dm, dc;
    check(gk_dcsr_upload(context(), &a, &da.d), "gk_dcsr_upload");
    check(gk_dcsr_upload(context(), &b, &db.d), "gk_dcsr_upload");
    check(gk_dcsr_upload(context(), &cm.csr, &dm.d), "gk_dcsr_upload");
    check(gk_masked_spgemm_cmask_dev(context(), da.d, db.d, dm.d, GK_CMASK_AUTO, &dc.d),
          "gk_masked_spgemm_cmask_dev");
    int64_t m = 0, n = 0, nz = 0;
    check(gk_dcsr_dims(dc.d, &m, &n, &nz), "gk_dcsr_dims");
    spMtx<T> R((size_t)m, (size_t)n, (size_t)nz);
    std::memcpy(R.matcode, C.matcode, sizeof(MM_typecode));
    check(gk_dcsr_download(context(), dc.d, R.Rst.data(), R.Col.data()), "gk_dcsr_download");
    check(gk_dcsr_download_val(context(), dc.d, R.Val.data()), "gk_dcsr_download_val");
    C = std::move(R);
}

}  // namespace detail
}  // namespace gk

// C<M> = A*B, accumulator chosen per row by the library (heuristic, or the
// learned selector when gk::options() asks for it).
template <typename T, typename U = T>
void mxmm_gpu(bool isParallel, const spMtx<T> &A, const spMtx<T> &B, const spMtx<U> &M, spMtx<T> &C) {
    (void)isParallel;
    gk::detail::masked_product(A, B, M, C);
}

// include/matrix_la.h:106-111, 562-574
template <typename T, typename U = T>
void mxmm_msa(bool isParallel, const spMtx<T> &A, const spMtx<T> &B, const spMtx<U> &M, spMtx<T> &C) {
    (void)isParallel;
    gk::detail::masked_product(A, B, M, C);
}

// include/matrix_la.h:87-92, 478-491
template <typename T, typename U = T>
void mxmm_mca(bool isParallel, const spMtx<T> &A, const spMtx<T> &B, const spMtx<U> &M, spMtx<T> &C) {
    (void)isParallel;
    gk::detail::masked_product(A, B, M, C);
}

// Returning form, same parameter types as include/matrix_la.h:80-84 (the
// value types of A and B must agree for the product to run).
template <typename T, typename U>
spMtx<T> mxmm_mca(bool isParallel, const spMtx<T> &A, const spMtx<U> &B, const spMtx<T> &M) {
    static_assert(std::is_same<T, U>::value, "A and B must have the same value type");
    spMtx<T> C;
    mxmm_mca(isParallel, A, B, M, C);
    return C;
}

// include/matrix_la.h:126-131, 733-934: C<not M> = A*B.  C gets M's shape
// and every (i,j) that some term A(i,k)*B(k,j) reaches and M does not hold,
// rows in column order, explicit zeros kept, int32 wrap-around.  Repeated
// mask columns count once (the reference's symbolic pass would subtract
// them twice and mis-size the row).
template <typename T, typename U = T>
void mxmm_msa_cmask(bool isParallel, const spMtx<T> &A, const spMtx<T> &B, const spMtx<U> &M, spMtx<T> &C) {
    (void)isParallel;
    gk::detail::complement_product(A, B, M, C);
}
