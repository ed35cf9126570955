// graphkokkos/algorithms.h -- the graph drivers of GraphKokkos's
// src/main.cpp in a header: triangle_counting_masked_sandia
// (src/main.cpp:69-90), k_truss (src/main.cpp:93-141) and the two unwired
// variants triangle_counting_vertex and triangle_counting_masked_lu
// (src/main.cpp:16-66) with their signatures, plus the full truss
// decomposition and the support product.
//
// Given one of this library's GPU products (mxmm_gpu, mxmm_msa, mxmm_mca)
// the drivers run the fused GPU entry points (gk_tc_count, gk_ktruss),
// which never materialise C; given any other mxmOp<int> they run the
// reference's loop around that product.  Both print the reference's
// "Time:" / "Triangles:" / "Iterations:" lines.
#pragma once

#include <chrono>
#include <cstdint>
#include <iostream>
#include <memory>
#include <vector>

#include "matrix_la.h"

namespace gk {
namespace detail {

inline bool is_gpu_product(mxmOp<int> op) {
    return op == static_cast<mxmOp<int>>(&mxmm_gpu<int, int>) ||
           op == static_cast<mxmOp<int>>(&mxmm_msa<int, int>) ||
           op == static_cast<mxmOp<int>>(&mxmm_mca<int, int>);
}

template <typename T>
bool rows_sorted(const spMtx<T> &A) {
    for (size_t i = 0; i < A.m; ++i)
        for (int p = A.Rst(i) + 1; p < A.Rst(i + 1); ++p)
            if (A.Col(p) <= A.Col(p - 1)) return false;
    return true;
}

// Sorted symmetric pattern without self loops whose values are all 1: the
// input the fused truss entry points take.
inline bool is_unit_undirected(const spMtx<int> &A) {
    if (A.m != A.n || !rows_sorted(A)) return false;
    for (size_t p = 0; p < A.nz; ++p)
        if (A.Val.size() >= A.nz && A.Val(p) != 1) return false;
    for (size_t i = 0; i < A.m; ++i)
        for (int p = A.Rst(i); p < A.Rst(i + 1); ++p) {
            const int j = A.Col(p);
            if ((size_t)j == i) return false;
            if (!std::binary_search(A.Col.data() + A.Rst(j), A.Col.data() + A.Rst(j + 1), (int)i)) return false;
        }
    return true;
}

inline long long ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace detail
}  // namespace gk

// Number of triangles: sum of C<L> = L*L with L the strictly lower part of A
// (src/main.cpp:69-90).  The count is 64-bit.
inline int64_t triangle_counting_masked_sandia(const spMtx<int> &A, mxmOp<int> matrixMult = mxmm_gpu<int, int>,
                                               bool isParallel = true) {
    const spMtx<int> L = extract_lower_triangle(A);
    int64_t num_of_tr = 0;
    const auto t0 = std::chrono::steady_clock::now();
    if (gk::detail::is_gpu_product(matrixMult) && gk::detail::rows_sorted(L)) {
        const gk_csr l = L.csr(false);
        uint64_t cnt = 0;
        gk::check(gk_tc_count(gk::context(), &l, &cnt, &gk::options()), "gk_tc_count");
        num_of_tr = (int64_t)cnt;
    } else {
        spMtx<int> C;
        matrixMult(isParallel, L, L, L, C);
        for (size_t p = 0; p < C.nz; ++p) num_of_tr += C.Val(p);
    }
    std::cout << "Time:       " << gk::detail::ms_since(t0) << " ms\n";
    std::cout << "Triangles:  " << num_of_tr << '\n';
    return num_of_tr;
}

// Triangles through every vertex: the int sum of each row of SQ<A> = A*A,
// shifted right once (src/main.cpp:16-39).  As in the reference the result
// is new int[A.m], which the caller deletes, and only "Time:" is printed.
// Given one of this library's products and sorted rows, gk_tc_vertex takes
// the sums on the device and SQ never reaches the host.
inline int *triangle_counting_vertex(const spMtx<int> &A, mxmOp<int> matrixMult = mxmm_gpu<int, int>,
                                     bool isParallel = true) {
    std::unique_ptr<int[]> nums_of_tr(new int[A.m]);
    const auto t0 = std::chrono::steady_clock::now();
    if (gk::detail::is_gpu_product(matrixMult) && A.m == A.n && gk::detail::rows_sorted(A)) {
        const gk_csr a = A.csr();
        gk::check(gk_tc_vertex(gk::context(), &a, nums_of_tr.get(), &gk::options()), "gk_tc_vertex");
    } else {
        spMtx<int> SQ;
        matrixMult(isParallel, A, A, A, SQ);
        for (size_t i = 0; i < A.m; ++i) {
            uint32_t s = 0;  // int sum, wrapping
            for (int j = SQ.Rst(i); j < SQ.Rst(i + 1); ++j) s += (uint32_t)SQ.Val(j);
            nums_of_tr[i] = (int)s >> 1;
        }
    }
    std::cout << "Time:       " << gk::detail::ms_since(t0) << '\n';
    return nums_of_tr.release();
}

// Number of triangles as sum(C<A> = L*U) / 2, L the strictly lower part of
// A and U its transpose (src/main.cpp:42-66).  The product is matrixMult.
inline int64_t triangle_counting_masked_lu(const spMtx<int> &A, mxmOp<int> matrixMult = mxmm_gpu<int, int>,
                                           bool isParallel = true) {
    const spMtx<int> L = extract_lower_triangle(A);
    const spMtx<int> U = transpose(L);
    spMtx<int> C;
    const auto t0 = std::chrono::steady_clock::now();
    matrixMult(isParallel, L, U, A, C);
    int64_t num_of_tr = 0;
    for (size_t p = 0; p < C.nz; ++p) num_of_tr += C.Val(p);
    num_of_tr >>= 1;
    std::cout << "Time:       " << gk::detail::ms_since(t0) << " ms\n";
    std::cout << "Triangles:  " << num_of_tr << '\n';
    return num_of_tr;
}

// The k-truss of A: repeat S<C> = C*C and keep the entries with S >= k-2
// until no entry is removed (src/main.cpp:93-141).  The result is A's
// surviving pattern with every value 1.
inline spMtx<int> k_truss(const spMtx<int> &A, int k, mxmOp<int> matrixMult = mxmm_gpu<int, int>,
                          bool isParallel = true) {
    const auto t0 = std::chrono::steady_clock::now();
    int64_t iterations = 0;
    spMtx<int> C;
    if (gk::detail::is_gpu_product(matrixMult) && k >= 2 && gk::detail::is_unit_undirected(A)) {
        const gk_csr a = A.csr(false);
        std::vector<int32_t> rst(A.m + 1), col(A.nz ? A.nz : 1);
        int64_t nz = 0;
        gk::check(gk_ktruss(gk::context(), &a, k, rst.data(), col.data(), &nz, &iterations, &gk::options()),
                  "gk_ktruss");
        C = spMtx<int>(A.m, A.n, (size_t)nz);
        std::memcpy(C.matcode, A.matcode, sizeof(MM_typecode));
        std::copy(rst.begin(), rst.end(), C.Rst.data());
        std::copy(col.begin(), col.begin() + nz, C.Col.data());
        std::fill(C.Val.data(), C.Val.data() + nz, 1);
    } else {
        C = A;
        for (size_t t = 0; t < A.m; ++t) {
            spMtx<int> S;
            matrixMult(isParallel, C, C, C, S);
            spMtx<int> R(C.m, C.n, C.nz);
            std::memcpy(R.matcode, C.matcode, sizeof(MM_typecode));
            size_t w = 0;
            for (size_t i = 0; i < C.m; ++i) {
                for (int p = S.Rst(i); p < S.Rst(i + 1); ++p)
                    if (S.Val(p) >= k - 2) { R.Col(w) = S.Col(p); R.Val(w++) = 1; }
                R.Rst(i + 1) = (int)w;
            }
            const bool unchanged = w == C.nz;
            R.resizeVals(w);
            C = std::move(R);
            iterations = (int64_t)t + 1;
            if (unchanged) break;
        }
    }
    std::cout << "Time:       " << gk::detail::ms_since(t0) << " ms\n";
    std::cout << "Iterations: " << iterations << '\n';
    return C;
}

// Truss number of every entry of the symmetric pattern A: the largest k
// whose k-truss holds the edge, 2 for edges in no triangle.
inline std::vector<uint32_t> truss_decomposition(const spMtx<int> &A, int32_t *kmax = nullptr) {
    const gk_csr a = A.csr(false);
    std::vector<uint32_t> truss(A.nz);
    int32_t km = 0;
    gk::check(gk_truss_decompose(gk::context(), &a, truss.data(), &km, &gk::options()), "gk_truss_decompose");
    if (kmax) *kmax = km;
    return truss;
}

// Triangles through every edge: S<A> = A*A on A's pattern (the "mxm" action
// of src/main.cpp:228-233 on a pattern graph).
inline std::vector<uint32_t> edge_support(const spMtx<int> &A) {
    const gk_csr a = A.csr(false);
    std::vector<uint32_t> s(A.nz);
    gk::check(gk_ktruss_support(gk::context(), &a, s.data(), &gk::options()), "gk_ktruss_support");
    return s;
}
