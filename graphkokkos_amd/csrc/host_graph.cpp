// Host-side graph construction for libgraphhip.so: seeded generators
// (R-MAT, Erdos-Renyi, Chung-Lu, triangulated grid), the cleaning pipeline
// (symmetrise, drop self loops, dedup, degree relabel, sorted int32 CSR) and
// the Matrix Market reader with GraphKokkos's semantics.
//
// Reference behaviour restated here (paths relative to the GraphKokkos tree):
//   read_mtx_to_crs          include/matrix.h:259-395
//   build_adjacency_matrix   include/matrix.h:643-660  (pattern, values 1)
//   extract_lower_triangle   include/matrix.h:698-726  (strict lower part)
//   build_symm_from_lower    include/matrix.h:662-696  (L merged with L^T)
// The reference's own generator (include/matrix_utils.h) is unseeded and is
// not reproduced; the generators below are this library's own.
#include <cmath>
#include <cstring>
#include <cctype>
#include <fstream>
#include <memory>
#include <numeric>
#include <sstream>

#include "gk_internal.h"

namespace gk {

static thread_local std::string g_err;
static int g_host_threads = 0;

void set_error(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}
void clear_error() { g_err.clear(); }
// Set on the workers of a batch build: each builds whole graphs on its own.
static thread_local bool t_serial = false;

void chung_lu_cdf(int64_t n, double exponent, std::vector<double> &cdf) {
    cdf.resize(n);
    const double p = -1.0 / (exponent - 1.0);
    double acc = 0;
    for (int64_t i = 0; i < n; ++i) { acc += std::pow(double(i + 1), p); cdf[i] = acc; }
    for (auto &x : cdf) x /= acc;
    cdf[n - 1] = 1.0;
}

int host_threads() {
    if (t_serial) return 1;
    if (g_host_threads > 0) return g_host_threads;
    unsigned h = std::thread::hardware_concurrency();
    return h ? (int)std::min(h, 64u) : 1;
}

}  // namespace gk

using namespace gk;

struct gk_hgraph {
    int64_t n = 0;
    std::vector<int32_t> arst, acol;   // symmetric pattern
    std::vector<int32_t> lrst, lcol;   // strictly lower part
    std::vector<int32_t> perm;         // new id of old vertex
};

struct gk_mtx {
    int64_t m = 0, n = 0, nz = 0;
    std::vector<int32_t> rst, col;
    std::vector<double> val;
    char matcode[4] = {' ', ' ', ' ', 'G'};
};

extern "C" const char *gk_last_error(void) { return gk::g_err.c_str(); }
extern "C" int gk_abi_version(void) { return GK_ABI_VERSION; }
extern "C" void gk_set_host_threads(int n) { gk::g_host_threads = n > 0 ? n : 0; }

namespace {

// Edge source: f(e, &u, &v) returns false for an edge to drop.
template <class EdgeFn>
int build_clean_graph(int64_t n, int64_t ne, EdgeFn &&edge, int relabel, gk_hgraph **out) {
    if (n <= 0 || n > INT32_MAX - 1) { set_error("vertex count %lld out of int32 range", (long long)n); return GK_ERANGE; }
    // 1. canonical (hi, lo) pairs, self loops dropped; bucket by hi.
    std::vector<int32_t> eu(ne), ev(ne);
    par_blocks(ne, [&](int64_t b, int64_t e, int) {
        for (int64_t i = b; i < e; ++i) {
            int64_t u, v;
            if (!edge(i, u, v) || u == v) { eu[i] = -1; ev[i] = -1; continue; }
            eu[i] = (int32_t)std::max(u, v);
            ev[i] = (int32_t)std::min(u, v);
        }
    });
    std::vector<int64_t> off(n + 1, 0);
    par_blocks(ne, [&](int64_t b, int64_t e, int) {
        for (int64_t i = b; i < e; ++i)
            if (eu[i] >= 0) __atomic_fetch_add(&off[eu[i] + 1], 1, __ATOMIC_RELAXED);
    });
    for (int64_t i = 0; i < n; ++i) off[i + 1] += off[i];
    const int64_t total = off[n];
    std::vector<int32_t> lo(total);
    {
        std::vector<int64_t> cur(off.begin(), off.end() - 1);
        par_blocks(ne, [&](int64_t b, int64_t e, int) {
            for (int64_t i = b; i < e; ++i)
                if (eu[i] >= 0) lo[__atomic_fetch_add(&cur[eu[i]], 1, __ATOMIC_RELAXED)] = ev[i];
        });
    }
    std::vector<int32_t>().swap(eu);
    std::vector<int32_t>().swap(ev);
    // 2. sort + unique every bucket.
    std::vector<int64_t> cnt(n, 0);
    par_dynamic(n, 1024, [&](int64_t r) {
        int32_t *b = lo.data() + off[r], *e = lo.data() + off[r + 1];
        std::sort(b, e);
        cnt[r] = std::unique(b, e) - b;
    });
    // 3. degrees in old labels.
    std::vector<int64_t> deg(n, 0);
    par_blocks(n, [&](int64_t b, int64_t e, int) {
        for (int64_t r = b; r < e; ++r) {
            __atomic_fetch_add(&deg[r], cnt[r], __ATOMIC_RELAXED);
            for (int64_t p = off[r]; p < off[r] + cnt[r]; ++p)
                __atomic_fetch_add(&deg[lo[p]], 1, __ATOMIC_RELAXED);
        }
    });
    int64_t nnzA = 0;
    for (int64_t r = 0; r < n; ++r) nnzA += deg[r];
    if (nnzA > INT32_MAX) { set_error("nnz %lld exceeds int32 CSR", (long long)nnzA); return GK_ERANGE; }
    // 4. relabel by (degree desc, id asc).
    auto g = std::make_unique<gk_hgraph>();
    g->n = n;
    g->perm.resize(n);
    if (relabel) {
        std::vector<int32_t> order(n);
        std::iota(order.begin(), order.end(), 0);
        std::sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
            return deg[x] != deg[y] ? deg[x] > deg[y] : x < y;
        });
        for (int64_t r = 0; r < n; ++r) g->perm[order[r]] = (int32_t)r;
    } else {
        std::iota(g->perm.begin(), g->perm.end(), 0);
    }
    // 5. symmetric A in new labels, rows sorted.
    g->arst.assign(n + 1, 0);
    for (int64_t v = 0; v < n; ++v) g->arst[g->perm[v] + 1] = (int32_t)deg[v];
    for (int64_t i = 0; i < n; ++i) g->arst[i + 1] += g->arst[i];
    g->acol.resize(nnzA);
    {
        std::vector<int32_t> cur(g->arst.begin(), g->arst.end() - 1);
        par_blocks(n, [&](int64_t b, int64_t e, int) {
            for (int64_t r = b; r < e; ++r) {
                int32_t a = g->perm[r];
                for (int64_t p = off[r]; p < off[r] + cnt[r]; ++p) {
                    int32_t c = g->perm[lo[p]];
                    g->acol[__atomic_fetch_add(&cur[a], 1, __ATOMIC_RELAXED)] = c;
                    g->acol[__atomic_fetch_add(&cur[c], 1, __ATOMIC_RELAXED)] = a;
                }
            }
        });
    }
    par_dynamic(n, 1024, [&](int64_t r) {
        std::sort(g->acol.begin() + g->arst[r], g->acol.begin() + g->arst[r + 1]);
    });
    // 6. strictly lower part.
    g->lrst.assign(n + 1, 0);
    par_blocks(n, [&](int64_t b, int64_t e, int) {
        for (int64_t r = b; r < e; ++r) {
            auto rb = g->acol.begin() + g->arst[r], re = g->acol.begin() + g->arst[r + 1];
            g->lrst[r + 1] = (int32_t)(std::lower_bound(rb, re, (int32_t)r) - rb);
        }
    });
    for (int64_t i = 0; i < n; ++i) g->lrst[i + 1] += g->lrst[i];
    g->lcol.resize(g->lrst[n]);
    par_blocks(n, [&](int64_t b, int64_t e, int) {
        for (int64_t r = b; r < e; ++r)
            std::copy(g->acol.begin() + g->arst[r], g->acol.begin() + g->arst[r] + (g->lrst[r + 1] - g->lrst[r]),
                      g->lcol.begin() + g->lrst[r]);
    });
    *out = g.release();
    return GK_OK;
}

}  // namespace

extern "C" int gk_hgraph_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                              int relabel, gk_hgraph **out) {
    GK_CHECK_ARG(out, "out is NULL");
    GK_CHECK_ARG(scale >= 1 && scale <= 30, "R-MAT scale %d out of range [1,30]", scale);
    GK_CHECK_ARG(edge_factor >= 1, "edge factor must be positive");
    GK_CHECK_ARG(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0, "R-MAT probabilities invalid");
    const int64_t n = int64_t(1) << scale;
    const int64_t ne = n * edge_factor;
    const double ab = a + b, abc = a + b + c;
    // Edge e descends `scale` quadrant levels; level l draws rng(seed, e, l).
    return build_clean_graph(n, ne, [&](int64_t e, int64_t &u, int64_t &v) {
        int64_t x = 0, y = 0;
        for (int l = 0; l < scale; ++l) {
            double r = u01(rng(seed, (uint64_t)e, (uint64_t)l));
            int q = r < a ? 0 : (r < ab ? 1 : (r < abc ? 2 : 3));
            x = (x << 1) | (q >> 1);
            y = (y << 1) | (q & 1);
        }
        u = x; v = y;
        return true;
    }, relabel, out);
}

extern "C" int gk_hgraph_er(int64_t n, double mean_degree, uint64_t seed, int relabel, gk_hgraph **out) {
    GK_CHECK_ARG(out, "out is NULL");
    GK_CHECK_ARG(n >= 2 && mean_degree > 0, "ER needs n >= 2 and a positive mean degree");
    // G(n, m) with m = n*d/2 uniform pairs (duplicates then removed).
    const int64_t ne = (int64_t)std::llround(n * mean_degree / 2.0);
    return build_clean_graph(n, ne, [&](int64_t e, int64_t &u, int64_t &v) {
        u = (int64_t)(u01(rng(seed, (uint64_t)e, 0)) * n);
        v = (int64_t)(u01(rng(seed, (uint64_t)e, 1)) * n);
        if (u >= n) u = n - 1;
        if (v >= n) v = n - 1;
        return true;
    }, relabel, out);
}

extern "C" int gk_hgraph_chung_lu(int64_t n, double exponent, double mean_degree, uint64_t seed, int relabel,
                                  gk_hgraph **out) {
    GK_CHECK_ARG(out, "out is NULL");
    GK_CHECK_ARG(n >= 2 && exponent > 1.0 && mean_degree > 0, "Chung-Lu needs n >= 2, exponent > 1, degree > 0");
    // Weight of vertex i is (i+1)^(-1/(exponent-1)); both endpoints of each of
    // the m = n*d/2 edges are drawn proportionally to weight (inverse CDF).
    std::vector<double> cdf;
    chung_lu_cdf(n, exponent, cdf);
    const int64_t ne = (int64_t)std::llround(n * mean_degree / 2.0);
    auto draw = [&](uint64_t r) {
        double x = u01(r);
        return (int64_t)(std::upper_bound(cdf.begin(), cdf.end(), x) - cdf.begin());
    };
    return build_clean_graph(n, ne, [&](int64_t e, int64_t &u, int64_t &v) {
        u = std::min<int64_t>(draw(rng(seed, (uint64_t)e, 0)), n - 1);
        v = std::min<int64_t>(draw(rng(seed, (uint64_t)e, 1)), n - 1);
        return true;
    }, relabel, out);
}

extern "C" int gk_hgraph_batch(int64_t ngraphs, const int32_t *kind, const int64_t *n, const double *exponent,
                               const double *mean_degree, const uint64_t *seed, int relabel, gk_hgraph **out,
                               int64_t *voff) {
    GK_CHECK_ARG(out && voff && kind && n && mean_degree && seed, "null argument");
    GK_CHECK_ARG(ngraphs >= 1, "ngraphs must be positive");
    for (int64_t g = 0; g < ngraphs; ++g) {
        GK_CHECK_ARG(kind[g] == 0 || kind[g] == 1, "graph %lld: kind must be 0 (ER) or 1 (Chung-Lu)", (long long)g);
        GK_CHECK_ARG(n[g] >= 2 && mean_degree[g] > 0, "graph %lld: needs n >= 2 and a positive degree", (long long)g);
        GK_CHECK_ARG(kind[g] == 0 || (exponent && exponent[g] > 1.0), "graph %lld: exponent must be > 1",
                     (long long)g);
    }
    voff[0] = 0;
    for (int64_t g = 0; g < ngraphs; ++g) voff[g + 1] = voff[g] + n[g];
    GK_CHECK_ARG(voff[ngraphs] < INT32_MAX, "total vertex count exceeds int32 CSR");
    // every worker builds whole graphs with the single-threaded pipeline
    std::vector<gk_hgraph *> parts(ngraphs, nullptr);
    std::vector<int> rcs(ngraphs, GK_OK);
    std::vector<std::string> errs(ngraphs);
    {
        std::atomic<int64_t> next{0};
        std::vector<std::thread> th;
        const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(host_threads(), ngraphs));
        for (int t = 0; t < nt; ++t)
            th.emplace_back([&] {
                t_serial = true;
                for (;;) {
                    const int64_t g = next.fetch_add(1);
                    if (g >= ngraphs) break;
                    rcs[g] = kind[g] == 0 ? gk_hgraph_er(n[g], mean_degree[g], seed[g], relabel, &parts[g])
                                          : gk_hgraph_chung_lu(n[g], exponent[g], mean_degree[g], seed[g], relabel,
                                                               &parts[g]);
                    if (rcs[g] != GK_OK) errs[g] = gk_last_error();
                }
                t_serial = false;
            });
        for (auto &x : th) x.join();
    }
    auto release = [&] { for (auto *q : parts) delete q; };
    for (int64_t g = 0; g < ngraphs; ++g)
        if (rcs[g] != GK_OK) {
            set_error("graph %lld: %s", (long long)g, errs[g].c_str());
            release();
            return rcs[g];
        }
    // block-diagonal union: graph g on vertices [voff[g], voff[g+1])
    int64_t nnzA = 0, nnzL = 0;
    for (auto *q : parts) { nnzA += (int64_t)q->acol.size(); nnzL += (int64_t)q->lcol.size(); }
    if (nnzA >= INT32_MAX) { release(); set_error("union nnz %lld exceeds int32 CSR", (long long)nnzA); return GK_ERANGE; }
    auto u = std::make_unique<gk_hgraph>();
    const int64_t N = voff[ngraphs];
    u->n = N;
    u->arst.assign(N + 1, 0);
    u->lrst.assign(N + 1, 0);
    u->acol.resize(nnzA);
    u->lcol.resize(nnzL);
    u->perm.resize(N);
    std::vector<int64_t> aoff(ngraphs + 1, 0), loff(ngraphs + 1, 0);
    for (int64_t g = 0; g < ngraphs; ++g) {
        aoff[g + 1] = aoff[g] + (int64_t)parts[g]->acol.size();
        loff[g + 1] = loff[g] + (int64_t)parts[g]->lcol.size();
    }
    par_dynamic(ngraphs, 16, [&](int64_t g) {
        const gk_hgraph *q = parts[g];
        const int32_t v0 = (int32_t)voff[g];
        for (int64_t i = 0; i < q->n; ++i) {
            u->arst[v0 + i + 1] = (int32_t)(aoff[g] + q->arst[i + 1]);
            u->lrst[v0 + i + 1] = (int32_t)(loff[g] + q->lrst[i + 1]);
            u->perm[v0 + i] = v0 + q->perm[i];
        }
        for (size_t p = 0; p < q->acol.size(); ++p) u->acol[aoff[g] + p] = v0 + q->acol[p];
        for (size_t p = 0; p < q->lcol.size(); ++p) u->lcol[loff[g] + p] = v0 + q->lcol[p];
    });
    release();
    *out = u.release();
    return GK_OK;
}

extern "C" int gk_hgraph_grid(int64_t rows, int64_t cols, int diagonal, int relabel, gk_hgraph **out) {
    GK_CHECK_ARG(out, "out is NULL");
    GK_CHECK_ARG(rows >= 1 && cols >= 1, "grid needs positive dimensions");
    // Edge slots per cell (r,c): right, down, and down-right when diagonal.
    const int64_t n = rows * cols;
    const int64_t per = diagonal ? 3 : 2;
    return build_clean_graph(n, n * per, [&](int64_t e, int64_t &u, int64_t &v) {
        int64_t cell = e / per, kind = e % per;
        int64_t r = cell / cols, c = cell % cols;
        u = cell;
        if (kind == 0) { if (c + 1 >= cols) return false; v = cell + 1; }
        else if (kind == 1) { if (r + 1 >= rows) return false; v = cell + cols; }
        else { if (r + 1 >= rows || c + 1 >= cols) return false; v = cell + cols + 1; }
        return true;
    }, relabel, out);
}

extern "C" int gk_hgraph_from_edges(int64_t n, int64_t ne, const int32_t *u, const int32_t *v, int relabel,
                                    gk_hgraph **out) {
    GK_CHECK_ARG(out && (ne == 0 || (u && v)), "null argument");
    for (int64_t e = 0; e < ne; ++e)
        GK_CHECK_ARG(u[e] >= 0 && u[e] < n && v[e] >= 0 && v[e] < n, "edge %lld out of range", (long long)e);
    return build_clean_graph(n, ne, [&](int64_t e, int64_t &x, int64_t &y) {
        x = u[e]; y = v[e];
        return true;
    }, relabel, out);
}

extern "C" int gk_hgraph_from_csr(const gk_csr *A, int relabel, gk_hgraph **out) {
    GK_CHECK_ARG(A && out && A->rst && (A->nz == 0 || A->col), "null argument");
    GK_CHECK_ARG(A->m == A->n, "graph matrix must be square");
    // Every entry with col < row, as extract_lower_triangle keeps on sorted
    // rows (include/matrix.h:707-712), then mirrored.
    std::vector<int32_t> eu, ev;
    for (int64_t i = 0; i < A->m; ++i)
        for (int64_t p = A->rst[i]; p < A->rst[i + 1]; ++p)
            if (A->col[p] < i) { eu.push_back((int32_t)i); ev.push_back(A->col[p]); }
    return gk_hgraph_from_edges(A->m, (int64_t)eu.size(), eu.data(), ev.data(), relabel, out);
}

extern "C" int gk_hgraph_view(const gk_hgraph *g, gk_csr *A, gk_csr *L) {
    GK_CHECK_ARG(g, "graph is NULL");
    if (A) *A = gk_csr{g->n, g->n, (int64_t)g->acol.size(), g->arst.data(), g->acol.data(), nullptr};
    if (L) *L = gk_csr{g->n, g->n, (int64_t)g->lcol.size(), g->lrst.data(), g->lcol.data(), nullptr};
    return GK_OK;
}

extern "C" int gk_hgraph_perm(const gk_hgraph *g, int32_t *new_of_old) {
    GK_CHECK_ARG(g && new_of_old, "null argument");
    std::copy(g->perm.begin(), g->perm.end(), new_of_old);
    return GK_OK;
}

extern "C" void gk_hgraph_free(gk_hgraph *g) { delete g; }

// ---- Matrix Market -------------------------------------------------------

namespace {
std::string lower(std::string s) {
    for (auto &ch : s) ch = (char)std::tolower((unsigned char)ch);
    return s;
}
}  // namespace

extern "C" int gk_mtx_read(const char *path, gk_mtx **out) {
    GK_CHECK_ARG(path && out, "null argument");
    std::ifstream in(path);
    if (!in.is_open()) { set_error("cannot open %s", path); return GK_EIO; }
    auto mt = std::make_unique<gk_mtx>();
    // Banner: "%%MatrixMarket matrix coordinate <type> <symmetry>", fields
    // compared case-insensitively (src/mmio.cpp:96-179).
    std::string line;
    if (!std::getline(in, line)) { set_error("%s: empty file", path); return GK_EIO; }
    std::istringstream bs(line);
    std::string banner, obj, fmt, type, sym;
    if (!(bs >> banner >> obj >> fmt >> type >> sym)) { set_error("%s: short banner", path); return GK_EIO; }
    if (banner != "%%MatrixMarket") { set_error("%s: no MatrixMarket banner", path); return GK_EIO; }
    obj = lower(obj); fmt = lower(fmt); type = lower(type); sym = lower(sym);
    if (obj != "matrix") { set_error("%s: unsupported object %s", path, obj.c_str()); return GK_EIO; }
    mt->matcode[0] = 'M';
    if (fmt == "coordinate") mt->matcode[1] = 'C';
    else if (fmt == "array") mt->matcode[1] = 'A';
    else { set_error("%s: unsupported format %s", path, fmt.c_str()); return GK_EIO; }
    if (type == "real") mt->matcode[2] = 'R';
    else if (type == "complex") mt->matcode[2] = 'C';
    else if (type == "pattern") mt->matcode[2] = 'P';
    else if (type == "integer") mt->matcode[2] = 'I';
    else { set_error("%s: unsupported type %s", path, type.c_str()); return GK_EIO; }
    if (sym == "general") mt->matcode[3] = 'G';
    else if (sym == "symmetric") mt->matcode[3] = 'S';
    else if (sym == "hermitian") mt->matcode[3] = 'H';
    else if (sym == "skew-symmetric") mt->matcode[3] = 'K';
    else { set_error("%s: unsupported symmetry %s", path, sym.c_str()); return GK_EIO; }
    // Complex and dense array files and non-square matrices are rejected
    // (include/matrix.h:282-289).
    if (mt->matcode[2] == 'C' || mt->matcode[1] == 'A') {
        set_error("%s: complex or array matrices are not supported", path);
        return GK_EIO;
    }
    // Size line: first line not starting with '%'.
    do {
        if (!std::getline(in, line)) { set_error("%s: missing size line", path); return GK_EIO; }
    } while (!line.empty() && line[0] == '%');
    long long M = 0, N = 0, NZ = 0;
    {
        std::istringstream ss(line);
        if (!(ss >> M >> N >> NZ)) { set_error("%s: bad size line", path); return GK_EIO; }
    }
    if (M != N) { set_error("%s: not a square matrix", path); return GK_EIO; }
    if (M <= 0 || M > INT32_MAX - 1 || NZ < 0) { set_error("%s: sizes out of range", path); return GK_ERANGE; }
    const bool pattern = mt->matcode[2] == 'P';
    const bool symm = mt->matcode[3] == 'S';
    std::vector<int32_t> ri, ci;
    std::vector<double> vi;
    ri.reserve(symm ? 2 * NZ : NZ);
    ci.reserve(symm ? 2 * NZ : NZ);
    vi.reserve(symm ? 2 * NZ : NZ);
    for (long long t = 0; t < NZ; ++t) {
        long long r, c;
        double x = 1.0;
        if (!(in >> r >> c)) { set_error("%s: entry %lld unreadable", path, t); return GK_EIO; }
        if (!pattern && !(in >> x)) { set_error("%s: value %lld unreadable", path, t); return GK_EIO; }
        if (r < 1 || r > M || c < 1 || c > N) { set_error("%s: entry %lld out of range", path, t); return GK_EIO; }
        if (r == c) continue;  // self loops dropped (include/matrix.h:325-328)
        ri.push_back((int32_t)(r - 1)); ci.push_back((int32_t)(c - 1)); vi.push_back(x);
        if (symm) { ri.push_back((int32_t)(c - 1)); ci.push_back((int32_t)(r - 1)); vi.push_back(x); }
    }
    if ((long long)ri.size() > INT32_MAX) { set_error("%s: nnz exceeds int32", path); return GK_ERANGE; }
    // Stable counting sort by row: column order within a row = file order
    // (include/matrix.h:370-386), duplicates kept.
    mt->m = M; mt->n = N; mt->nz = (int64_t)ri.size();
    mt->rst.assign(M + 1, 0);
    for (auto r : ri) ++mt->rst[r + 1];
    for (long long i = 0; i < M; ++i) mt->rst[i + 1] += mt->rst[i];
    mt->col.resize(mt->nz);
    mt->val.resize(mt->nz);
    std::vector<int32_t> cur(mt->rst.begin(), mt->rst.end() - 1);
    for (size_t t = 0; t < ri.size(); ++t) {
        int32_t p = cur[ri[t]]++;
        mt->col[p] = ci[t];
        mt->val[p] = vi[t];
    }
    *out = mt.release();
    return GK_OK;
}

extern "C" int gk_mtx_view(const gk_mtx *mt, gk_csr *pattern, const double **val, char matcode[4]) {
    GK_CHECK_ARG(mt, "mtx is NULL");
    if (pattern) *pattern = gk_csr{mt->m, mt->n, mt->nz, mt->rst.data(), mt->col.data(), nullptr};
    if (val) *val = mt->val.data();
    if (matcode) std::memcpy(matcode, mt->matcode, 4);
    return GK_OK;
}

extern "C" void gk_mtx_free(gk_mtx *mt) { delete mt; }

// ---- the other on-disk formats --------------------------------------------
// Readers with the layouts of include/matrix.h:396-533 and writers after
// :209-257, into / from the same gk_mtx as the Matrix Market reader.
//   crs    text: "m nz matcode", then nz columns, nz values, m+1 row pointers
//   bin    4 matcode bytes, m n nz as 64-bit sizes, int32 row pointers [m+1],
//          int32 columns [nz], int32 values [nz] (the reference's writer puts
//          the row pointers in all three arrays, :220-222; here each array
//          gets its own data)
//   graph  text: "m nz", then one line of 0-based columns per row
//   rmat   text: "m n nz", then nz pairs "x y" (0-based); both directions
//          are stored, rows sorted and deduplicated (the reference keeps
//          insertion order and repeats, :488-533)
namespace {

bool csr_consistent(const gk_mtx &mt, const char *path) {
    if (mt.rst.empty() || mt.rst[0] != 0 || mt.rst[mt.m] != mt.nz) {
        set_error("%s: row pointers must run from 0 to nz", path);
        return false;
    }
    for (int64_t i = 0; i < mt.m; ++i)
        if (mt.rst[i + 1] < mt.rst[i]) { set_error("%s: row pointers decrease at row %lld", path, (long long)i); return false; }
    for (int64_t p = 0; p < mt.nz; ++p)
        if (mt.col[p] < 0 || mt.col[p] >= mt.n) { set_error("%s: column %lld out of range", path, (long long)p); return false; }
    return true;
}

int read_crs(const char *path, gk_mtx &mt) {
    std::ifstream in(path);
    if (!in.is_open()) { set_error("cannot open %s", path); return GK_EIO; }
    long long m = 0, nz = 0;
    std::string code;
    if (!(in >> m >> nz >> code)) { set_error("%s: bad crs header", path); return GK_EIO; }
    if (m < 0 || m > INT32_MAX - 1 || nz < 0 || nz > INT32_MAX) { set_error("%s: sizes out of range", path); return GK_ERANGE; }
    // The reference leaves n unset for this format; a square matrix is assumed.
    mt.m = mt.n = m;
    mt.nz = nz;
    std::memcpy(mt.matcode, "MCIG", 4);
    for (size_t x = 0; x < 4 && x < code.size(); ++x) mt.matcode[x] = code[x];
    mt.col.resize(nz);
    mt.val.resize(nz);
    mt.rst.resize(m + 1);
    for (long long p = 0; p < nz; ++p)
        if (!(in >> mt.col[p])) { set_error("%s: column %lld unreadable", path, p); return GK_EIO; }
    for (long long p = 0; p < nz; ++p)
        if (!(in >> mt.val[p])) { set_error("%s: value %lld unreadable", path, p); return GK_EIO; }
    for (long long i = 0; i <= m; ++i)
        if (!(in >> mt.rst[i])) { set_error("%s: row pointer %lld unreadable", path, i); return GK_EIO; }
    return csr_consistent(mt, path) ? GK_OK : GK_EIO;
}

int read_bin(const char *path, gk_mtx &mt) {
    std::unique_ptr<FILE, int (*)(FILE *)> fp(std::fopen(path, "rb"), std::fclose);
    if (!fp) { set_error("cannot open %s", path); return GK_EIO; }
    uint64_t mnz[3];
    if (std::fread(mt.matcode, 1, 4, fp.get()) != 4 || std::fread(mnz, sizeof(uint64_t), 3, fp.get()) != 3) {
        set_error("%s: short bin header", path);
        return GK_EIO;
    }
    if (mnz[0] > INT32_MAX - 1 || mnz[1] > INT32_MAX - 1 || mnz[2] > INT32_MAX) {
        set_error("%s: sizes out of range", path);
        return GK_ERANGE;
    }
    mt.m = (int64_t)mnz[0];
    mt.n = (int64_t)mnz[1];
    mt.nz = (int64_t)mnz[2];
    mt.rst.resize(mt.m + 1);
    mt.col.resize(mt.nz);
    std::vector<int32_t> v(mt.nz);
    if (std::fread(mt.rst.data(), 4, mt.m + 1, fp.get()) != (size_t)(mt.m + 1) ||
        std::fread(mt.col.data(), 4, mt.nz, fp.get()) != (size_t)mt.nz ||
        std::fread(v.data(), 4, mt.nz, fp.get()) != (size_t)mt.nz) {
        set_error("%s: short bin body", path);
        return GK_EIO;
    }
    mt.val.assign(v.begin(), v.end());
    return csr_consistent(mt, path) ? GK_OK : GK_EIO;
}

int read_graph(const char *path, gk_mtx &mt) {
    std::ifstream in(path);
    if (!in.is_open()) { set_error("cannot open %s", path); return GK_EIO; }
    long long m = 0, nz = 0;
    std::string line;
    if (!std::getline(in, line)) { set_error("%s: empty file", path); return GK_EIO; }
    {
        std::istringstream hs(line);
        if (!(hs >> m >> nz)) { set_error("%s: bad graph header", path); return GK_EIO; }
    }
    if (m < 0 || m > INT32_MAX - 1 || nz < 0 || nz > INT32_MAX) { set_error("%s: sizes out of range", path); return GK_ERANGE; }
    mt.m = mt.n = m;
    std::memcpy(mt.matcode, "MCIG", 4);
    mt.rst.assign(m + 1, 0);
    mt.col.clear();
    mt.col.reserve(nz);
    for (long long i = 0; i < m; ++i) {
        if (!std::getline(in, line)) line.clear();  // missing trailing rows are empty
        std::istringstream ls(line);
        long long c;
        while (ls >> c) {
            if (c < 0 || c >= m) { set_error("%s: row %lld: column %lld out of range", path, i, c); return GK_EIO; }
            mt.col.push_back((int32_t)c);
        }
        if ((long long)mt.col.size() > nz) { set_error("%s: more than nz = %lld columns", path, nz); return GK_EIO; }
        mt.rst[i + 1] = (int32_t)mt.col.size();
    }
    if ((long long)mt.col.size() != nz) { set_error("%s: %zu columns for nz = %lld", path, mt.col.size(), nz); return GK_EIO; }
    mt.nz = nz;
    mt.val.assign(nz, 1.0);
    return GK_OK;
}

int read_rmat(const char *path, gk_mtx &mt) {
    std::ifstream in(path);
    if (!in.is_open()) { set_error("cannot open %s", path); return GK_EIO; }
    long long m = 0, n = 0, ne = 0;
    if (!(in >> m >> n >> ne)) { set_error("%s: bad rmat header", path); return GK_EIO; }
    if (m < 0 || m > INT32_MAX - 1 || n != m || ne < 0 || 2 * ne > INT32_MAX) {
        set_error("%s: sizes out of range (a square matrix is required)", path);
        return GK_ERANGE;
    }
    std::vector<std::vector<int32_t>> rows(m);
    for (long long e = 0; e < ne; ++e) {
        long long x, y;
        if (!(in >> x >> y)) { set_error("%s: pair %lld unreadable", path, e); return GK_EIO; }
        if (x < 0 || x >= m || y < 0 || y >= m) { set_error("%s: pair %lld out of range", path, e); return GK_EIO; }
        rows[x].push_back((int32_t)y);
        rows[y].push_back((int32_t)x);
    }
    mt.m = mt.n = m;
    std::memcpy(mt.matcode, "MCIG", 4);
    mt.rst.assign(m + 1, 0);
    mt.col.clear();
    for (long long i = 0; i < m; ++i) {
        auto &r = rows[i];
        std::sort(r.begin(), r.end());
        r.erase(std::unique(r.begin(), r.end()), r.end());
        mt.col.insert(mt.col.end(), r.begin(), r.end());
        mt.rst[i + 1] = (int32_t)mt.col.size();
    }
    mt.nz = (int64_t)mt.col.size();
    mt.val.assign(mt.nz, 1.0);
    return GK_OK;
}

}  // namespace

extern "C" int gk_matrix_read(const char *path, const char *format, gk_mtx **out) {
    GK_CHECK_ARG(path && format && out, "null argument");
    const std::string f = format;
    if (f == "mtx") return gk_mtx_read(path, out);
    auto mt = std::make_unique<gk_mtx>();
    int rc;
    if (f == "crs") rc = read_crs(path, *mt);
    else if (f == "bin") rc = read_bin(path, *mt);
    else if (f == "graph") rc = read_graph(path, *mt);
    else if (f == "rmat") rc = read_rmat(path, *mt);
    else { set_error("%s: unknown format %s", path, format); return GK_EINVAL; }
    if (rc != GK_OK) return rc;
    *out = mt.release();
    return GK_OK;
}

extern "C" int gk_matrix_write(const char *path, const char *format, const gk_csr *A, const int32_t *val,
                               const char matcode[4]) {
    GK_CHECK_ARG(path && format && A && A->rst && (A->col || A->nz == 0), "null argument");
    const std::string f = format;
    char code[4];
    if (matcode) std::memcpy(code, matcode, 4);
    else std::memcpy(code, val ? "MCIG" : "MCPG", 4);
    if (f == "bin") {
        std::unique_ptr<FILE, int (*)(FILE *)> fp(std::fopen(path, "wb"), std::fclose);
        if (!fp) { set_error("cannot create %s", path); return GK_EIO; }
        const uint64_t mnz[3] = {(uint64_t)A->m, (uint64_t)A->n, (uint64_t)A->nz};
        std::vector<int32_t> ones;
        if (!val) ones.assign(A->nz, 1);
        const int32_t *v = val ? val : ones.data();
        bool ok = std::fwrite(code, 1, 4, fp.get()) == 4 && std::fwrite(mnz, sizeof(uint64_t), 3, fp.get()) == 3 &&
                  std::fwrite(A->rst, 4, A->m + 1, fp.get()) == (size_t)(A->m + 1) &&
                  std::fwrite(A->col, 4, A->nz, fp.get()) == (size_t)A->nz &&
                  std::fwrite(v, 4, A->nz, fp.get()) == (size_t)A->nz;
        if (!ok) { set_error("%s: write failed", path); return GK_EIO; }
        return GK_OK;
    }
    std::ofstream o(path);
    if (!o.is_open()) { set_error("cannot create %s", path); return GK_EIO; }
    if (f == "mtx") {
        // general coordinate; integer values, or pattern without values
        // (write_crs_to_mtx, include/matrix.h:228-257)
        o << "%%MatrixMarket matrix coordinate " << (val ? "integer" : "pattern") << " general\n";
        o << A->m << ' ' << A->n << ' ' << A->nz << '\n';
        for (int64_t i = 0; i < A->m; ++i)
            for (int32_t p = A->rst[i]; p < A->rst[i + 1]; ++p) {
                o << i + 1 << ' ' << A->col[p] + 1;
                if (val) o << ' ' << val[p];
                o << '\n';
            }
    } else if (f == "crs") {
        o << A->m << ' ' << A->nz << ' ' << std::string(code, 4) << '\n';
        for (int64_t p = 0; p < A->nz; ++p) o << A->col[p] << (p + 1 < A->nz ? ' ' : '\n');
        for (int64_t p = 0; p < A->nz; ++p) o << (val ? val[p] : 1) << (p + 1 < A->nz ? ' ' : '\n');
        for (int64_t i = 0; i <= A->m; ++i) o << A->rst[i] << (i < A->m ? ' ' : '\n');
    } else {
        set_error("%s: the writers are bin, mtx and crs (not %s)", path, format);
        return GK_EINVAL;
    }
    if (!o) { set_error("%s: write failed", path); return GK_EIO; }
    return GK_OK;
}
