// graphkokkos/betweenness_centrality.h -- drop-in for GraphKokkos's
// include/betweenness_centrality.h: batched Brandes betweenness on the GPU
// through gk_bc_batch_f64.
//
// betweenness_centrality_batch keeps the reference's signature and meaning
// (include/betweenness_centrality.h:406-506): the sources are vertices
// 0 .. min(n, batchSize)-1, bc[v] = sum over those sources s of the
// dependency delta_s(v) on all targets (unnormalised, without the 1/2 of
// undirected betweenness), and the result comes back as float.  Path
// counts and dependencies are FP64 on the device; the overload
// betweenness_centrality_batch_f64 takes any source list and returns the
// FP64 values.
#pragma once

#include <algorithm>
#include <cstdint>
#include <vector>

#include "matrix.h"

inline std::vector<double> betweenness_centrality_batch_f64(const spMtx<int> &A, const std::vector<int> &sources) {
    if (A.m != A.n) throw "non-square matrix; BC is only for square matrices";
    const gk_csr a = A.csr(false);
    std::vector<double> bc(A.m);
    gk::check(gk_bc_batch_f64(gk::context(), &a, sources.data(), (int64_t)sources.size(), bc.data(),
                              &gk::options()),
              "gk_bc_batch_f64");
    return bc;
}

inline std::vector<float> betweenness_centrality_batch(bool isParallel, const spMtx<int> &A, size_t batchSize) {
    (void)isParallel;
    if (A.m != A.n) throw "non-square matrix; BC is only for square matrices";
    const size_t ns = std::min(A.n, batchSize);
    std::vector<int> sources(ns);
    for (size_t s = 0; s < ns; ++s) sources[s] = (int)s;
    const std::vector<double> bc = betweenness_centrality_batch_f64(A, sources);
    return std::vector<float>(bc.begin(), bc.end());
}
