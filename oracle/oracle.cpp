// ORACLE -- test infrastructure only.  This file is the CPU restatement of
// GraphKokkos's algorithms that the tests, the smoke check and the CPU
// baseline leg of bench.py compare libgraphhip.so against.  It is never
// linked into, loaded by or called from the product library.
//
// Each function restates a reference routine (paths relative to the
// GraphKokkos source tree; the reference itself needs Kokkos, which is not
// installed, so it is not built here):
//
//   or_mxmm_msa   _mxmm_msa_parallel    include/matrix_la.h:576-635
//                 _mxmm_msa_sequential  include/matrix_la.h:696-731
//   or_mxmm_mca   _mxmm_mca_parallel    include/matrix_la.h:493-528
//                 _mxmm_mca_sequential  include/matrix_la.h:530-560
//   or_tc_sandia  triangle_counting_masked_sandia  src/main.cpp:69-90
//   or_ktruss     k_truss                           src/main.cpp:93-141
//   or_truss_decompose  k_truss looped over k (every edge's largest k)
//   or_truss_peel       independent queue-based peeling (cross-check of
//                       or_truss_decompose, fast enough for large graphs)
//   or_mxmm_cmask _mxmm_msa_cmask_sequential  include/matrix_la.h:853-934
//                 and its parallel twin :745-851 (C<not M> = A*B:
//                 symbolic row counts, then values)
//   or_bc_batch   betweenness_centrality_batch  include/betweenness_centrality.h:406-506
//                 with bc_forward_step (:313-401, OpenMP twin :248-309) and
//                 bc_backward_step (:182-242, OpenMP twin :128-178)
//
// How each is pinned (tests/test_oracle.py): closed forms (K_n, triangulated
// grid, paths, stars, bipartite graphs), brute force on tiny graphs
// (dense A*A, textbook Brandes by BFS), networkx/scipy golden fixtures
// committed under tests/golden, and cross-checks between the independent
// restatements (MSA vs MCA, sequential vs parallel, decompose vs peel).
#include <omp.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct Csr {
    int64_t m, n, nz;
    const int32_t *rst, *col, *val;  // val == nullptr: all ones
    int32_t v(int64_t p) const { return val ? val[p] : 1; }
};

// Row ranges of the reference's nnz-balanced static partition
// (include/matrix_la.h:590-596): team t owns rows [wd[t], wd[t+1]).  The
// reference overruns wd when A.nz == 0 (every row passes the test); the
// restatement stops at the last team instead.
std::vector<int64_t> work_distribution(const Csr &A, int64_t teams) {
    std::vector<int64_t> wd(teams + 1, 0);
    int64_t team_i = 1;
    for (int64_t i = 0; i <= A.m && team_i <= teams; ++i)
        if ((int64_t)A.rst[i] >= (int64_t)A.nz * team_i / teams) wd[team_i++] = i;
    return wd;
}

// Parallel MSA: per-thread dense accumulator of B.n entries; per row, zero
// the mask columns, scatter every B-row entry, gather at the mask columns.
void msa_parallel(const Csr &A, const Csr &B, const Csr &M, int32_t *cval) {
    const int64_t teams = omp_get_max_threads();
    std::vector<int64_t> wd = work_distribution(A, teams);
#pragma omp parallel num_threads(teams)
    {
        std::vector<int32_t> acc(B.n, 0);
        int64_t t = omp_get_thread_num();
        for (int64_t i = wd[t]; i < wd[t + 1]; ++i) {
            for (int64_t j = M.rst[i]; j < M.rst[i + 1]; ++j) acc[M.col[j]] = 0;
            for (int64_t p = A.rst[i]; p < A.rst[i + 1]; ++p) {
                int32_t k = A.col[p];
                uint32_t a = (uint32_t)A.v(p);
                for (int64_t q = B.rst[k]; q < B.rst[k + 1]; ++q)  // int wrap as unsigned
                    acc[B.col[q]] = (int32_t)((uint32_t)acc[B.col[q]] + a * (uint32_t)B.v(q));
            }
            for (int64_t j = M.rst[i]; j < M.rst[i + 1]; ++j) cval[j] = acc[M.col[j]];
        }
    }
}

// Sequential MSA with the three states UNALLOWED / ALLOWED / SET.  As in the
// reference, a mask column that stays ALLOWED (no contribution) returns the
// accumulator's stale value from an earlier row; `faithful == 0` returns 0
// there instead, which is what the parallel MSA and both MCAs return.
void msa_sequential(const Csr &A, const Csr &B, const Csr &M, int32_t *cval, bool faithful) {
    enum : char { UNALLOWED = 0, ALLOWED = 1, SET = 2 };
    std::vector<char> state(B.n, UNALLOWED);
    std::vector<int32_t> value(B.n, 0);
    for (int64_t i = 0; i < A.m; ++i) {
        for (int64_t j = M.rst[i]; j < M.rst[i + 1]; ++j) state[M.col[j]] = ALLOWED;
        for (int64_t p = A.rst[i]; p < A.rst[i + 1]; ++p) {
            int32_t k = A.col[p];
            uint32_t a = (uint32_t)A.v(p);
            for (int64_t q = B.rst[k]; q < B.rst[k + 1]; ++q) {
                int32_t c = B.col[q];
                uint32_t prod = a * (uint32_t)B.v(q);
                if (state[c] == ALLOWED) { state[c] = SET; value[c] = (int32_t)prod; }
                else if (state[c] == SET) value[c] = (int32_t)((uint32_t)value[c] + prod);
            }
        }
        for (int64_t j = M.rst[i]; j < M.rst[i + 1]; ++j) {
            int32_t c = M.col[j];
            cval[j] = (faithful || state[c] == SET) ? value[c] : 0;
            state[c] = UNALLOWED;
        }
    }
}

// MCA: accumulator indexed by mask position; each A entry (i,k) merge-walks
// B row k against mask row i (needs sorted columns).
void mca_row(const Csr &A, const Csr &B, const Csr &M, int64_t i, int32_t *acc, int32_t *cval) {
    int64_t mlen = M.rst[i + 1] - M.rst[i];
    for (int64_t p = A.rst[i]; p < A.rst[i + 1]; ++p) {
        int32_t k = A.col[p];
        int64_t b = B.rst[k], be = B.rst[k + 1];
        uint32_t a = (uint32_t)A.v(p);
        int64_t mp = M.rst[i];
        for (int64_t j = 0; j < mlen; ++j, ++mp) {
            while (b < be && B.col[b] < M.col[mp]) ++b;
            if (b < be && B.col[b] == M.col[mp]) acc[j] = (int32_t)((uint32_t)acc[j] + a * (uint32_t)B.v(b));
        }
    }
    for (int64_t j = 0; j < mlen; ++j) { cval[M.rst[i] + j] = acc[j]; acc[j] = 0; }
}

void mca(const Csr &A, const Csr &B, const Csr &M, int32_t *cval, bool parallel) {
    int64_t len = 0;
    for (int64_t i = 0; i < A.m; ++i) len = std::max<int64_t>(len, M.rst[i + 1] - M.rst[i]);
    if (!parallel) {
        std::vector<int32_t> acc(len, 0);
        for (int64_t i = 0; i < A.m; ++i) mca_row(A, B, M, i, acc.data(), cval);
        return;
    }
#pragma omp parallel
    {
        std::vector<int32_t> acc(len, 0);
#pragma omp for schedule(dynamic)
        for (int64_t i = 0; i < A.m; ++i) mca_row(A, B, M, i, acc.data(), cval);
    }
}

void product(int accum, int parallel, const Csr &A, const Csr &B, const Csr &M, int32_t *cval) {
    if (accum == 1) mca(A, B, M, cval, parallel != 0);
    else if (parallel) msa_parallel(A, B, M, cval);
    else msa_sequential(A, B, M, cval, accum == 2);
}

// Compacted CSR over the surviving entries of a symmetric pattern, with the
// id of each surviving entry in the original pattern.
struct Sub {
    int64_t m;
    std::vector<int32_t> rst, col, id;
    Csr csr() const { return Csr{m, m, (int64_t)col.size(), rst.data(), col.data(), nullptr}; }
};

// One run of the reference's k_truss loop (src/main.cpp:104-131) on `C`:
// Tmp<C> = C*C, keep entries with Tmp >= k-2 (values reset to 1), repeat
// until an iteration removes nothing.  Returns the iteration count as the
// reference reports it.
int64_t ktruss_loop(Sub &C, int k, int accum, int parallel) {
    std::vector<int32_t> s;
    for (int64_t t = 0; t < C.m || C.m == 0; ++t) {
        s.assign(C.col.size(), 0);
        Csr c = C.csr();
        product(accum, parallel, c, c, c, s.data());
        int64_t pos = 0;
        Sub T;
        T.m = C.m;
        T.rst.assign(C.m + 1, 0);
        for (int64_t i = 0; i < C.m; ++i) {
            for (int64_t j = C.rst[i]; j < C.rst[i + 1]; ++j)
                if (s[j] >= k - 2) { T.col.push_back(C.col[j]); T.id.push_back(C.id[j]); ++pos; }
            T.rst[i + 1] = (int32_t)pos;
        }
        if (T.col.size() == C.col.size()) return t + 1;
        C = std::move(T);
        if (C.m == 0) break;
    }
    return 0;
}

Sub full_sub(const Csr &A) {
    Sub S;
    S.m = A.m;
    S.rst.assign(A.rst, A.rst + A.m + 1);
    S.col.assign(A.col, A.col + A.nz);
    S.id.resize(A.nz);
    for (int64_t p = 0; p < A.nz; ++p) S.id[p] = (int32_t)p;
    return S;
}

}  // namespace

extern "C" {

int or_num_threads(void) { return omp_get_max_threads(); }
void or_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }

// accum: 0 = MSA, 1 = MCA, 2 = MSA sequential with the reference's stale
// read kept (parallel must be 0 for 2).
int or_mxmm(int accum, int parallel,
            int64_t am, int64_t an, int64_t anz, const int32_t *arst, const int32_t *acol, const int32_t *aval,
            int64_t bm, int64_t bn, int64_t bnz, const int32_t *brst, const int32_t *bcol, const int32_t *bval,
            int64_t mm, int64_t mn, int64_t mnz, const int32_t *mrst, const int32_t *mcol,
            int32_t *cval) {
    if (an != bm || am != mm || bn != mn) return -1;
    Csr A{am, an, anz, arst, acol, aval}, B{bm, bn, bnz, brst, bcol, bval}, M{mm, mn, mnz, mrst, mcol, nullptr};
    product(accum, parallel, A, B, M, cval);
    return 0;
}

// C<not M> = A*B by the reference's complement-mask MSA: pass 1 counts, per
// row, the distinct columns some A(i,k)*B(k,j) reaches minus those in M's
// row; pass 2 accumulates them (mask columns UNALLOWED) and writes them in
// column order, explicit zeros kept.  Called with col == NULL it fills rst
// only and returns nz; then again with col and val.  `parallel` splits the
// rows over OpenMP threads, each with its own accumulator, as the
// reference's parallel variant does (include/matrix_la.h:745-851); the
// rows' results do not depend on it.
int64_t or_mxmm_cmask(int parallel, int64_t am, int64_t an, int64_t anz, const int32_t *arst, const int32_t *acol,
                      const int32_t *aval, int64_t bm, int64_t bn, int64_t bnz, const int32_t *brst,
                      const int32_t *bcol, const int32_t *bval, int64_t mm, int64_t mn, int64_t mnz,
                      const int32_t *mrst, const int32_t *mcol, int32_t *rst, int32_t *col, int32_t *val) {
    if (an != bm || am != mm || bn != mn) return -1;
    Csr A{am, an, anz, arst, acol, aval}, B{bm, bn, bnz, brst, bcol, bval};
    (void)mnz;
    enum { UNALLOWED = 0, ALLOWED = 1, SET = 2 };
    if (!col) {
        std::vector<int64_t> cnt(A.m + 1, 0);
#pragma omp parallel if (parallel)
        {
            std::vector<uint8_t> state(bn, UNALLOWED);
            std::vector<int32_t> changed;
#pragma omp for schedule(dynamic, 64)
            for (int64_t i = 0; i < A.m; ++i) {
                int64_t row_nz = 0;
                for (int64_t t = A.rst[i]; t < A.rst[i + 1]; ++t) {
                    const int32_t k = A.col[t];
                    for (int64_t j = B.rst[k]; j < B.rst[k + 1]; ++j)
                        if (state[B.col[j]] == UNALLOWED) {
                            state[B.col[j]] = ALLOWED;
                            changed.push_back(B.col[j]);
                            ++row_nz;
                        }
                }
                for (int64_t j = mrst[i]; j < mrst[i + 1]; ++j)
                    if (state[mcol[j]] == ALLOWED) --row_nz;
                cnt[i] = row_nz;
                for (int32_t c : changed) state[c] = UNALLOWED;
                changed.clear();
            }
        }
        rst[0] = 0;
        for (int64_t i = 0; i < A.m; ++i) rst[i + 1] = (int32_t)(rst[i] + cnt[i]);
        return rst[A.m];
    }
#pragma omp parallel if (parallel)
    {
        std::vector<uint8_t> state(bn, ALLOWED);
        std::vector<uint32_t> value(bn, 0);
        std::vector<int32_t> changed;
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < A.m; ++i) {
            for (int64_t j = mrst[i]; j < mrst[i + 1]; ++j) state[mcol[j]] = UNALLOWED;
            for (int64_t t = A.rst[i]; t < A.rst[i + 1]; ++t) {
                const int32_t k = A.col[t];
                const uint32_t a = (uint32_t)A.v(t);
                for (int64_t j = B.rst[k]; j < B.rst[k + 1]; ++j) {
                    const int32_t c = B.col[j];
                    if (state[c] == ALLOWED) {
                        state[c] = SET;
                        changed.push_back(c);
                        value[c] = a * (uint32_t)B.v(j);
                    } else if (state[c] == SET) {
                        value[c] += a * (uint32_t)B.v(j);
                    }
                }
            }
            std::sort(changed.begin(), changed.end());
            int64_t pos = rst[i];
            for (int32_t c : changed) {
                col[pos] = c;
                val[pos++] = (int32_t)value[c];
                value[c] = 0;
                state[c] = ALLOWED;
            }
            changed.clear();
            for (int64_t j = mrst[i]; j < mrst[i + 1]; ++j) state[mcol[j]] = ALLOWED;
        }
    }
    return rst[A.m];
}

// triangle_counting_masked_sandia: C<L> = L*L with the chosen accumulator,
// then the int64 sum of C's values.
int64_t or_tc_sandia(int accum, int parallel, int64_t m, int64_t nz, const int32_t *rst, const int32_t *col) {
    Csr L{m, m, nz, rst, col, nullptr};
    std::vector<int32_t> c(nz, 0);
    product(accum, parallel, L, L, L, c.data());
    int64_t sum = 0;
#pragma omp parallel for reduction(+ : sum)
    for (int64_t p = 0; p < nz; ++p) sum += c[p];
    return sum;
}

// Supports S<A> = A*A at A's pattern (the "mxm" action, src/main.cpp:228-233).
int or_support(int accum, int parallel, int64_t m, int64_t nz, const int32_t *rst, const int32_t *col, int32_t *s) {
    Csr A{m, m, nz, rst, col, nullptr};
    product(accum, parallel, A, A, A, s);
    return 0;
}

// k_truss for one k.  keep[p] = 1 if entry p of A survives.  Returns the
// reference's iteration count.
int64_t or_ktruss(int accum, int parallel, int k, int64_t m, int64_t nz, const int32_t *rst, const int32_t *col,
                  uint8_t *keep) {
    Csr A{m, m, nz, rst, col, nullptr};
    Sub C = full_sub(A);
    int64_t it = ktruss_loop(C, k, accum, parallel);
    std::memset(keep, 0, nz);
    for (int32_t id : C.id) keep[id] = 1;
    return it;
}

// Truss decomposition by the reference loop: C_2 = A; C_k = k_truss(C_{k-1}, k);
// entries of C_{k-1} missing from C_k get truss number k-1.
int or_truss_decompose(int accum, int parallel, int64_t m, int64_t nz, const int32_t *rst, const int32_t *col,
                       uint32_t *truss, int64_t *total_iterations) {
    Csr A{m, m, nz, rst, col, nullptr};
    Sub C = full_sub(A);
    std::vector<uint8_t> alive(nz, 1);
    int64_t iters = 0;
    int k = 3;
    while (!C.col.empty()) {
        iters += ktruss_loop(C, k, accum, parallel);
        std::vector<uint8_t> now(nz, 0);
        for (int32_t id : C.id) now[id] = 1;
        for (int64_t p = 0; p < nz; ++p)
            if (alive[p] && !now[p]) { truss[p] = (uint32_t)(k - 1); alive[p] = 0; }
        ++k;
    }
    if (total_iterations) *total_iterations = iters;
    int kmax = 2;
    for (int64_t p = 0; p < nz; ++p) kmax = std::max<int>(kmax, (int)truss[p]);
    return nz ? kmax : 0;
}

// Independent truss decomposition: supports by sorted intersection, then a
// FIFO of edges whose support fell below k-2; removing edge (u,v) destroys
// every triangle (u,v,w) whose other two edges are still present and
// decrements those two edges once.
int or_truss_peel(int64_t m, int64_t nz, const int32_t *rst, const int32_t *col, uint32_t *truss) {
    // Undirected edge id per entry: id of (min,max) = its position in the
    // strict-lower scan order; mirror entries share it.
    std::vector<int64_t> eid(nz, -1);
    int64_t ne = 0;
    std::vector<int32_t> eu, ev;
    for (int64_t i = 0; i < m; ++i)
        for (int64_t p = rst[i]; p < rst[i + 1]; ++p)
            if (col[p] < i) { eid[p] = ne++; eu.push_back((int32_t)i); ev.push_back(col[p]); }
    auto find = [&](int32_t r, int32_t c) -> int64_t {
        const int32_t *b = col + rst[r], *e = col + rst[r + 1];
        const int32_t *x = std::lower_bound(b, e, c);
        return (x != e && *x == c) ? (x - col) : -1;
    };
    for (int64_t i = 0; i < m; ++i)
        for (int64_t p = rst[i]; p < rst[i + 1]; ++p)
            if (col[p] > i) eid[p] = eid[find(col[p], (int32_t)i)];
    // Common neighbours of u and v (both rows sorted).  Rows of similar
    // length are merged; when one row is 16x longer or more, every entry of
    // the shorter row is looked up in the longer one by a galloping search
    // that resumes where the previous lookup ended, so a hub costs about
    // min(deg) * log(max deg / min deg) instead of its whole row.  visit(a, b)
    // receives the positions of each common column in u's and v's rows.
    auto common = [&](int32_t u, int32_t v, auto &&visit) {
        int64_t a = rst[u], ae = rst[u + 1], b = rst[v], be = rst[v + 1];
        bool swapped = (ae - a) > (be - b);
        if (swapped) { std::swap(a, b); std::swap(ae, be); }
        if ((be - b) < 16 * (ae - a)) {
            while (a < ae && b < be) {
                if (col[a] < col[b]) ++a;
                else if (col[a] > col[b]) ++b;
                else {
                    if (swapped) visit(b, a); else visit(a, b);
                    ++a; ++b;
                }
            }
            return;
        }
        for (; a < ae && b < be; ++a) {
            const int32_t c = col[a];
            int64_t step = 1, lo = b;
            while (lo + step < be && col[lo + step] < c) { lo += step; step <<= 1; }
            const int32_t *x = std::lower_bound(col + lo, col + std::min(be, lo + step + 1), c);
            b = x - col;
            if (b < be && col[b] == c) {
                if (swapped) visit(b, a); else visit(a, b);
                ++b;
            }
        }
    };
    std::vector<int32_t> sup(ne, 0);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t e = 0; e < ne; ++e) {
        int32_t s = 0;
        common(eu[e], ev[e], [&](int64_t, int64_t) { ++s; });
        sup[e] = s;
    }
    std::vector<uint8_t> alive(ne, 1), queued(ne, 0);
    std::vector<uint32_t> tr(ne, 2);
    int64_t left = ne;
    std::vector<int64_t> q;
    for (int k = 3; left > 0; ++k) {
        q.clear();
        for (int64_t e = 0; e < ne; ++e)
            if (alive[e] && !queued[e] && sup[e] < k - 2) { q.push_back(e); queued[e] = 1; }
        for (size_t h = 0; h < q.size(); ++h) {
            int64_t e = q[h];
            common(eu[e], ev[e], [&](int64_t a, int64_t b) {
                int64_t e1 = eid[a], e2 = eid[b];
                if (alive[e1] && alive[e2]) {
                    for (int64_t f : {e1, e2}) {
                        --sup[f];
                        if (!queued[f] && sup[f] < k - 2) { queued[f] = 1; q.push_back(f); }
                    }
                }
            });
            alive[e] = 0;
            tr[e] = (uint32_t)(k - 1);
            --left;
        }
    }
    int kmax = ne ? 2 : 0;
    for (int64_t p = 0; p < nz; ++p) {
        truss[p] = eid[p] >= 0 ? tr[eid[p]] : 0;
        kmax = std::max<int>(kmax, (int)truss[p]);
    }
    return kmax;
}

// Batched Brandes over an explicit source list, processed in batches of
// `batch` sources.  Per batch, with n = number of sources in it:
//   Numsp(src[j], j) = 1;  Front = (rows src of A)^T  (betweenness_centrality.h:444-447)
//   forward:  Sigmas[d] = Front;  Numsp += Front;
//             Front(i,j) = sum_k AT(i,k) Front(k,j) where Numsp(i,j) == 0
//   backward: Bcu = 1, Nspinv = 1/Numsp;  for level k = d-1 .. 1:
//             W = Nspinv .* Bcu on the pattern of Sigmas[k];
//             Bcu(i,j) += (A*W)(i,j) * Numsp(i,j) on the pattern of Sigmas[k-1]
//   bc(i) += sum_j Bcu(i,j) - n
// fp64 = 1 keeps sigma and delta in double; fp64 = 0 follows the reference's
// int32 sigma / float delta (results then carry float rounding).
int or_bc_batch(int fp64, int64_t m, int64_t nz, const int32_t *rst, const int32_t *col, const int32_t *src,
                int64_t nsrc, int64_t batch, double *bc) {
    // AT by counting sort (transpose, include/matrix_la.h:155-179).
    std::vector<int32_t> trst(m + 1, 0), tcol(nz);
    for (int64_t p = 0; p < nz; ++p) ++trst[col[p] + 1];
    for (int64_t i = 0; i < m; ++i) trst[i + 1] += trst[i];
    {
        std::vector<int32_t> cur(trst.begin(), trst.end() - 1);
        for (int64_t i = 0; i < m; ++i)
            for (int64_t p = rst[i]; p < rst[i + 1]; ++p) tcol[cur[col[p]]++] = (int32_t)i;
    }
    std::fill(bc, bc + m, 0.0);
    for (int64_t b0 = 0; b0 < nsrc; b0 += batch) {
        const int64_t n = std::min(batch, nsrc - b0);
        const int32_t *S = src + b0;
        // Frontier as per-row sorted (source, sigma) lists; a level is a
        // vector of such rows.
        struct Level { std::vector<int64_t> rst; std::vector<int32_t> col; std::vector<double> val; };
        std::vector<Level> sig;
        std::vector<double> numsp(m * n, 0.0);
        std::vector<int32_t> numsp32(fp64 ? 0 : m * n, 0);
        for (int64_t j = 0; j < n; ++j) { numsp[S[j] * n + j] = 1; if (!fp64) numsp32[S[j] * n + j] = 1; }
        Level F;
        {
            // transpose(A.extractRows(sources)): entry (v, j) for v in row S[j].
            std::vector<std::vector<std::pair<int32_t, double>>> rows(m);
            for (int64_t j = 0; j < n; ++j)
                for (int64_t p = rst[S[j]]; p < rst[S[j] + 1]; ++p) rows[col[p]].push_back({(int32_t)j, 1.0});
            F.rst.assign(m + 1, 0);
            for (int64_t i = 0; i < m; ++i) {
                for (auto &e : rows[i]) { F.col.push_back(e.first); F.val.push_back(e.second); }
                F.rst[i + 1] = (int64_t)F.col.size();
            }
        }
        do {
            sig.push_back(F);
            // Numsp += Front  ("InitNumsp", :321-324)
            for (int64_t i = 0; i < m; ++i)
                for (int64_t p = F.rst[i]; p < F.rst[i + 1]; ++p) {
                    numsp[i * n + F.col[p]] += F.val[p];
                    if (!fp64) numsp32[i * n + F.col[p]] += (int32_t)F.val[p];
                }
            // Front' = (AT * Front) .* not(Numsp)  ("MSpGEMMSymbolic/Numeric", :342-399)
            std::vector<std::vector<std::pair<int32_t, double>>> out(m);
#pragma omp parallel
            {
                std::vector<double> acc(n, 0.0);
                std::vector<int32_t> acc32(n, 0);
#pragma omp for schedule(dynamic, 256)
                for (int64_t i = 0; i < m; ++i) {
                    for (int64_t t = trst[i]; t < trst[i + 1]; ++t) {
                        int32_t k = tcol[t];
                        for (int64_t p = F.rst[k]; p < F.rst[k + 1]; ++p) {
                            if (fp64) acc[F.col[p]] += F.val[p];
                            else acc32[F.col[p]] += (int32_t)F.val[p];
                        }
                    }
                    for (int64_t j = 0; j < n; ++j) {
                        double a = fp64 ? acc[j] : (double)acc32[j];
                        bool unreached = fp64 ? numsp[i * n + j] == 0 : numsp32[i * n + j] == 0;
                        if (unreached && a != 0) out[i].push_back({(int32_t)j, a});
                        acc[j] = 0; acc32[j] = 0;
                    }
                }
            }
            Level G;
            G.rst.assign(m + 1, 0);
            for (int64_t i = 0; i < m; ++i) {
                for (auto &e : out[i]) { G.col.push_back(e.first); G.val.push_back(e.second); }
                G.rst[i + 1] = (int64_t)G.col.size();
            }
            F = std::move(G);
        } while (!F.col.empty());
        const int64_t d = (int64_t)sig.size();
        if (fp64) {
            std::vector<double> bcu(m * n, 1.0), w;
            for (int64_t k = d - 1; k > 0; --k) {
                const Level &Fk = sig[k], &Nx = sig[k - 1];
                w.assign(Fk.col.size(), 0.0);
                for (int64_t i = 0; i < m; ++i)
                    for (int64_t p = Fk.rst[i]; p < Fk.rst[i + 1]; ++p)
                        w[p] = (1.0 / numsp[i * n + Fk.col[p]]) * bcu[i * n + Fk.col[p]];
#pragma omp parallel
                {
                    std::vector<double> acc(n, 0.0);
#pragma omp for schedule(dynamic, 256)
                    for (int64_t i = 0; i < m; ++i) {
                        if (Nx.rst[i] == Nx.rst[i + 1]) continue;
                        for (int64_t p = Nx.rst[i]; p < Nx.rst[i + 1]; ++p) acc[Nx.col[p]] = 0;
                        for (int64_t t = rst[i]; t < rst[i + 1]; ++t) {
                            int32_t kk = col[t];
                            for (int64_t p = Fk.rst[kk]; p < Fk.rst[kk + 1]; ++p) acc[Fk.col[p]] += w[p];
                        }
                        for (int64_t p = Nx.rst[i]; p < Nx.rst[i + 1]; ++p) {
                            int32_t j = Nx.col[p];
                            bcu[i * n + j] += acc[j] * numsp[i * n + j];
                        }
                    }
                }
            }
            for (int64_t i = 0; i < m; ++i) {
                double s = 0;
                for (int64_t j = 0; j < n; ++j) s += bcu[i * n + j];
                bc[i] += s - (double)n;
            }
        } else {
            std::vector<float> bcu(m * n, 1.0f), w, numspd(m * n), nspinv(m * n);
            for (int64_t x = 0; x < m * n; ++x) { numspd[x] = (float)numsp32[x]; nspinv[x] = 1.0f / numspd[x]; }
            for (int64_t k = d - 1; k > 0; --k) {
                const Level &Fk = sig[k], &Nx = sig[k - 1];
                w.assign(Fk.col.size(), 0.0f);
                for (int64_t i = 0; i < m; ++i)
                    for (int64_t p = Fk.rst[i]; p < Fk.rst[i + 1]; ++p)
                        w[p] = nspinv[i * n + Fk.col[p]] * bcu[i * n + Fk.col[p]];
#pragma omp parallel
                {
                    std::vector<float> acc(n, 0.0f);
#pragma omp for schedule(dynamic, 256)
                    for (int64_t i = 0; i < m; ++i) {
                        if (Nx.rst[i] == Nx.rst[i + 1]) continue;
                        for (int64_t p = Nx.rst[i]; p < Nx.rst[i + 1]; ++p) acc[Nx.col[p]] = 0;
                        for (int64_t t = rst[i]; t < rst[i + 1]; ++t) {
                            int32_t kk = col[t];
                            for (int64_t p = Fk.rst[kk]; p < Fk.rst[kk + 1]; ++p) acc[Fk.col[p]] += w[p];
                        }
                        for (int64_t p = Nx.rst[i]; p < Nx.rst[i + 1]; ++p) {
                            int32_t j = Nx.col[p];
                            bcu[i * n + j] += acc[j] * numspd[i * n + j];
                        }
                    }
                }
            }
            for (int64_t i = 0; i < m; ++i) {
                float s = 0;
                for (int64_t j = 0; j < n; ++j) s += bcu[i * n + j];
                bc[i] += (double)(s - (float)n);
            }
        }
    }
    return 0;
}

}  // extern "C"
