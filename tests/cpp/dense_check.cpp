// Host-only checks of the drop-in container helpers that never reach the GPU:
// denseMtx<ValT> (include/matrix.h:548-630 of GraphKokkos), the free
// deep_copy (:536-546), convertType (:632-641) and extract_upper_triangle
// (:728-755).  Reads a CSR as "m n nz rst... col... val..." from stdin and
// writes the results as JSON lines for tests/test_dropin_host.py,
// which compares them with numpy.
#include <cstdio>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "graphkokkos/matrix.h"

namespace {

template <class V>
void put(const char *name, const V &v, size_t n) {
    std::cout << "{\"" << name << "\": [";
    for (size_t i = 0; i < n; ++i) std::cout << (i ? "," : "") << v[i];
    std::cout << "]}\n";
}

template <class T>
void put_csr(const char *name, const spMtx<T> &A) {
    std::cout << "{\"" << name << "\": {\"m\": " << A.m << ", \"n\": " << A.n << ", \"nz\": " << A.nz
              << ", \"matcode\": \"" << std::string(A.matcode, 4) << "\", \"rst\": [";
    for (size_t i = 0; i <= A.m; ++i) std::cout << (i ? "," : "") << A.Rst(i);
    std::cout << "], \"col\": [";
    for (size_t i = 0; i < A.nz; ++i) std::cout << (i ? "," : "") << A.Col(i);
    std::cout << "], \"val\": [";
    for (size_t i = 0; i < A.nz; ++i) std::cout << (i ? "," : "") << A.Val(i);
    std::cout << "]}}\n";
}

}  // namespace

int main() {
    size_t m, n, nz;
    std::cin >> m >> n >> nz;
    spMtx<int> A(m, n, nz);
    for (size_t i = 0; i <= m; ++i) std::cin >> A.Rst(i);
    for (size_t i = 0; i < nz; ++i) std::cin >> A.Col(i);
    for (size_t i = 0; i < nz; ++i) std::cin >> A.Val(i);
    A.matcode[0] = 'M'; A.matcode[1] = 'C'; A.matcode[2] = 'I'; A.matcode[3] = 'G';

    // constructed from a sparse matrix: row-major scatter into zeros
    denseMtx<double> D(A);
    put("dense_from_sparse", D.Val, D.m * D.n);

    // sized constructor: zeros
    denseMtx<int> Z(3, 4);
    put("zeros", Z.Val, Z.m * Z.n);

    // copies share the values (the reference's View member)
    denseMtx<double> S(D);
    S.Val(0) += 100.0;  // m, n >= 1 in every case the test feeds
    std::cout << "{\"copy_shares\": " << (D.Val(0) == S.Val(0) ? 1 : 0) << "}\n";

    // assignment from a sparse matrix clears what was there first
    denseMtx<int> R(m, n);
    for (size_t i = 0; i < m * n; ++i) R.Val(i) = 7;
    R = A;
    put("assign_from_sparse", R.Val, R.m * R.n);
    std::cout << "{\"equal_after_assign\": " << (R == denseMtx<int>(A) ? 1 : 0) << "}\n";

    // resize keeps the first values of the flat array, zero-fills the rest
    denseMtx<int> G(A);
    G.resize(m + 1, n);
    put("resized", G.Val, G.m * G.n);

    // move leaves the source empty
    denseMtx<int> M(std::move(G));
    std::cout << "{\"moved\": [" << M.m << "," << M.n << "," << G.Val.size() << "]}\n";

    // free deep_copy: an independent copy; views of dst's old arrays keep them
    spMtx<int> dst(1, 1, 1);
    spMtx<int> alias = dst;
    deep_copy(dst, A);
    if (nz) dst.Val(0) += 1000;
    std::cout << "{\"deep_copy_independent\": " << (!nz || A.Val(0) + 1000 == dst.Val(0) ? 1 : 0)
              << ", \"alias_kept\": " << (alias.Rst.size() == 2 && alias.Col.size() == 1 ? 1 : 0) << "}\n";
    put_csr("deep_copy", dst);

    // convertType: same pattern, values cast
    spMtx<double> F = convertType<double>(A);
    put_csr("convert", F);

    put_csr("upper", extract_upper_triangle(A));
    put_csr("lower", extract_lower_triangle(A));
    return 0;
}
