// Drives the C++ drop-in headers the way GraphKokkos's src/main.cpp drives
// its own (load .mtx -> adjacency -> symmetrise -> product / TC / k-truss /
// BC through mxmOp<int> pointers) and writes every result as JSON for the
// Python parity test to compare with the oracle.
//
//   dropin_check <graph.mtx> <out.json> <k> <bc_batch> [loaded]
//
// With "loaded" only the products on the pattern as loaded (rows in file
// order, duplicates kept) and BC run: files whose rows are not sorted have
// no well-defined symmetric pattern for the truss drivers.
#include <algorithm>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "graphkokkos/algorithms.h"
#include "graphkokkos/betweenness_centrality.h"

namespace {

// A product that is not one of the library's own pointers: the drivers then
// run the reference's loop around it.
void wrapped_product(bool par, const spMtx<int> &A, const spMtx<int> &B, const spMtx<int> &M, spMtx<int> &C) {
    mxmm_gpu<int, int>(par, A, B, M, C);
}

template <class V>
void put_array(std::ofstream &f, const char *name, const V &v, bool last = false) {
    f << '"' << name << "\": [";
    for (size_t i = 0; i < v.size(); ++i) f << (i ? "," : "") << v[i];
    f << "]" << (last ? "\n" : ",\n");
}

template <class T>
std::vector<long long> view_vec(const gk::View<T> &v, size_t n) {
    return std::vector<long long>(v.data(), v.data() + n);
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s graph.mtx out.json k bc_batch [loaded]\n", argv[0]);
        return 2;
    }
    gk::ScopeGuard guard(0);
    const int k = std::stoi(argv[3]);
    const size_t batch = (size_t)std::stoll(argv[4]);
    const bool loaded_only = argc > 5 && std::string(argv[5]) == "loaded";

    spMtx<int> gr(argv[1], "mtx");
    const spMtx<int> A = build_adjacency_matrix(gr);
    const spMtx<int> S = build_symm_from_lower(extract_lower_triangle(A));

    // the loaded matrix as the reader left it (rows in file order)
    std::ofstream f(argv[2]);
    f << "{\n";
    f << "\"m\": " << A.m << ",\n\"nz_loaded\": " << A.nz << ",\n";
    put_array(f, "loaded_rst", view_vec(A.Rst, A.m + 1));
    put_array(f, "loaded_col", view_vec(A.Col, A.nz));

    // C<A> = A*A on the loaded (possibly unsorted, duplicated) pattern and
    // C<S> = S*S on the symmetric one, through mxmOp<int> pointers
    mxmOp<int> msa = mxmm_msa<int>;
    mxmOp<int> mca = mxmm_mca<int>;
    spMtx<int> C1;
    msa(true, A, A, A, C1);
    put_array(f, "mxm_loaded", view_vec(C1.Val, C1.nz));
    // C<not A> = A*A: the pairs two hops apart that are not neighbours
    spMtx<int> N1;
    mxmm_msa_cmask(true, A, A, A, N1);
    put_array(f, "cmask_rst", view_vec(N1.Rst, N1.m + 1));
    put_array(f, "cmask_col", view_vec(N1.Col, N1.nz));
    put_array(f, "cmask_val", view_vec(N1.Val, N1.nz));
    if (!loaded_only) {
        spMtx<int> C2;
        mca(false, S, S, S, C2);
        put_array(f, "sym_rst", view_vec(S.Rst, S.m + 1));
        put_array(f, "sym_col", view_vec(S.Col, S.nz));
        put_array(f, "mxm_sym", view_vec(C2.Val, C2.nz));
        const spMtx<int> C3 = mxmm_mca(true, S, S, S);
        f << "\"mxm_return_form_equal\": " << (C3 == C2 ? 1 : 0) << ",\n";

        const int64_t tc_fused = triangle_counting_masked_sandia(S, msa, true);
        const int64_t tc_loop = triangle_counting_masked_sandia(S, wrapped_product, true);
        f << "\"tc_fused\": " << tc_fused << ",\n\"tc_loop\": " << tc_loop << ",\n";

        // the unwired drivers: triangles per vertex and sum(C<A> = L*U) / 2
        int *tv = triangle_counting_vertex(S, msa, true);
        int *tv_loop = triangle_counting_vertex(S, wrapped_product, true);
        put_array(f, "tc_vertex", std::vector<int>(tv, tv + S.m));
        f << "\"tc_vertex_loop_equal\": " << (std::equal(tv, tv + S.m, tv_loop) ? 1 : 0) << ",\n";
        delete[] tv;
        delete[] tv_loop;
        f << "\"tc_lu\": " << triangle_counting_masked_lu(S, msa, true) << ",\n";
        f << "\"tc_lu_loop\": " << triangle_counting_masked_lu(S, wrapped_product, true) << ",\n";

        const spMtx<int> K1 = k_truss(S, k, mca, true);
        const spMtx<int> K2 = k_truss(S, k, wrapped_product, true);
        put_array(f, "ktruss_rst", view_vec(K1.Rst, K1.m + 1));
        put_array(f, "ktruss_col", view_vec(K1.Col, K1.nz));
        f << "\"ktruss_loop_equal\": " << (K1 == K2 ? 1 : 0) << ",\n";

        int32_t kmax = 0;
        const std::vector<uint32_t> truss = truss_decomposition(S, &kmax);
        put_array(f, "truss", truss);
        f << "\"kmax\": " << kmax << ",\n";
        put_array(f, "support", edge_support(S));
    }

    const std::vector<float> bc = betweenness_centrality_batch(true, gr, batch);
    f.precision(9);
    put_array(f, "bc", bc);
    bool threw = false;
    try {
        spMtx<int> rect(3, 4, 0);
        (void)betweenness_centrality_batch(true, rect, 2);
    } catch (const char *) {
        threw = true;
    }
    f << "\"bc_nonsquare_throws\": " << (threw ? 1 : 0) << "\n}\n";
    return f ? 0 : 1;
}
