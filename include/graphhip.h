/*
 * graphhip.h -- C ABI of libgraphhip.so: masked sparse-sparse products and
 * the triangle-counting, k-truss and batched Brandes betweenness drivers on
 * one AMD Instinct MI355X (gfx950).
 *
 * Every entry point takes plain pointers and sizes, returns 0 on success and
 * a negative GK_E* code on failure; gk_last_error() then names the failing
 * check (thread-local string).  Host-pointer entry points copy their inputs
 * to HBM, run, and copy results back; the *_dev entry points work on graphs
 * already resident in HBM (gk_dcsr) and are the ones that are timed.
 *
 * Reference interfaces each entry point replaces (paths relative to the
 * GraphKokkos source tree):
 *   gk_masked_spgemm_i32   mxmm_msa / mxmm_mca       include/matrix_la.h:87-123, 478-491, 562-574
 *   gk_tc_count            triangle_counting_masked_sandia   src/main.cpp:69-90
 *   gk_ktruss_support      "mxm" action  S<A>=A*A            src/main.cpp:228-233
 *   gk_ktruss              k_truss                           src/main.cpp:93-141
 *   gk_truss_decompose     k_truss looped over k=3..kmax     src/main.cpp:93-141
 *   gk_bc_batch_f64        betweenness_centrality_batch      include/betweenness_centrality.h:406-506
 *   gk_mtx_read            spMtx(filename,"mtx") / read_mtx_to_crs  include/matrix.h:44-61, 259-395
 *   gk_hgraph_*            build_adjacency_matrix / extract_lower_triangle /
 *                          build_symm_from_lower            include/matrix.h:643-726
 */
#ifndef GRAPHHIP_H
#define GRAPHHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GK_ABI_VERSION 1

/* ---- status codes ------------------------------------------------------ */
#define GK_OK          0
#define GK_EINVAL     -1   /* bad argument (null pointer, shape mismatch, unsorted row) */
#define GK_ENOMEM     -2   /* device or host allocation failed */
#define GK_EHIP       -3   /* a HIP runtime call failed */
#define GK_EIO        -4   /* file could not be opened or parsed */
#define GK_ERANGE     -5   /* size beyond what int32 CSR indices can address */
#define GK_ESTATE     -6   /* context / handle misuse */

const char *gk_last_error(void);
int gk_abi_version(void);

/* ---- CSR views (borrowed, host memory) ---------------------------------
 * Same arrays as spMtx<ValT> (include/matrix.h:36-42): Rst[m+1], Col[nz],
 * Val[nz].  val == NULL means a pattern matrix with every value 1 (what
 * build_adjacency_matrix produces, include/matrix.h:643-660).              */
typedef struct gk_csr {
    int64_t m, n, nz;
    const int32_t *rst;
    const int32_t *col;
    const int32_t *val;
} gk_csr;

/* ---- context ------------------------------------------------------------
 * Counterpart of Kokkos::initialize (src/main.cpp:358): binds a device,
 * creates the stream every kernel of this context runs on and owns all
 * device buffers.  One host thread per context; calls are stream-ordered
 * and not re-entrant for one context.                                      */
typedef struct gk_ctx gk_ctx;
int  gk_init(int device, gk_ctx **out);
void gk_finalize(gk_ctx *ctx);
int  gk_sync(gk_ctx *ctx);
/* Host thread count used by graph construction (0 = hardware threads). */
void gk_set_host_threads(int nthreads);

/* ---- accumulator selection / options ------------------------------------ */
enum gk_accum {
    GK_ACC_AUTO        = -1, /* decided per row by the selector                */
    GK_ACC_MSA_LDS     = 0,  /* mask row hashed into LDS, dense slot counters  */
    GK_ACC_MCA_WAVE    = 1,  /* sorted-list intersection, ballot/popcount      */
    GK_ACC_HASH_GLOBAL = 2,  /* open-addressing table in HBM scratch           */
    GK_ACC_DENSE_GLOBAL= 3,  /* dense column-indexed marks in HBM scratch      */
    GK_NUM_ACC         = 4
};
/* Kernels reported in gk_timing: the four accumulators (0-3), the MSA-LDS
 * kernel for long mask rows (4) and the truss-peeling support update (5). */
#define GK_KERNEL_MSA_LONG 4
#define GK_KERNEL_PEEL     5
#define GK_NUM_KERNELS     6
enum gk_selector {
    GK_SEL_HEURISTIC = 0,   /* fixed thresholds (the fallback)            */
    GK_SEL_LEARNED   = 1,   /* MLP cost model, needs gk_model_set/train   */
    GK_SEL_FORCED    = 2    /* every row uses opts.forced_accum           */
};

typedef struct gk_opts {
    int32_t struct_size;    /* sizeof(gk_opts); 0 or a NULL opts means defaults */
    int32_t selector;       /* enum gk_selector                           */
    int32_t forced_accum;   /* enum gk_accum, used with GK_SEL_FORCED     */
    int32_t mca_max_mask;   /* heuristic: mask rows up to this use MCA; -1 (default) =
                             * the product's mode: 48 for triangle counting, 16 for
                             * support and value products                            */
    int32_t lds_max_mask;   /* heuristic: mask rows up to this use MSA-LDS*/
    int32_t dense_min_mask; /* heuristic: mask rows from this use dense-global */
    int32_t split_flop;     /* rows above this many B entries are split   */
    int32_t block_threads;  /* workgroup size for the row kernels         */
    int32_t verbose;
    int32_t small_max_mask; /* MSA-LDS rows up to this (<= 128) use the short-row kernel (default 32) */
    int32_t blocks_small;   /* persistent workgroups per CU, short-row MSA-LDS kernel */
    int32_t blocks_long;    /* ... long-row MSA-LDS kernel (0: 16 for triangle
                             * counting, 8 for support and value products)       */
    int32_t blocks_mca;     /* ... MCA-wave kernel (0: 16 for triangle counting, 8 for
                             * support and value products; the batching step, whose
                             * support product is mostly short rows, 32)         */
    int32_t blocks_global;  /* ... hash-global and dense-global kernels              */
    int32_t batch_blocks;   /* batching: peeling workgroups per CU                   */
    int32_t batch_order;    /* batching: 0 = graphs largest first, 1 = given order   */
    int32_t hub_window;     /* triangle counting: B rows below this id (<= 32768) are
                             * also kept as bitmaps; a row is read as bitmap words
                             * (window kernels) or single bits (MCA-wave) where that
                             * reads less than the row; 0 = off, default 32768       */
    int32_t learned_margin; /* learned selector: it leaves the fixed-threshold choice of
                             * a row only where the network predicts that choice to
                             * cost more than 2^(margin/1000) times its best one;
                             * 0 = plain argmin, default 250                         */
    int32_t hub_probe;      /* triangle counting with hub bitmaps: a window task tests
                             * single bits of a hub row where hub_probe times the mask
                             * row's columns below it is less than the row's words and
                             * entries (a test is a gathered word, the others are
                             * streamed); 0 = off, -1 (default) = 2                  */
    int32_t selector_labels;/* gk_selector_samples: 1 = concurrent (MSA-LDS and MCA-wave
                             * timed together on alternating halves of the rows, the
                             * engine's real schedule), 0 = each accumulator forced
                             * alone; default 1                                      */
} gk_opts;
void gk_opts_default(gk_opts *o);

/* Per-call timings (device time from HIP events on the context stream) and
 * algorithmic HBM bytes of the product kernels, summed over the launches of
 * the call (k-truss: every round; BC: every level).                         */
typedef struct gk_timing {
    double  total_ms;                 /* whole entry point, device side          */
    double  kernel_ms;                /* accumulator kernels, all launches       */
    double  select_ms;                /* features + selection + binning          */
    int64_t launches;                 /* product launches                        */
    int64_t iterations;               /* k-truss rounds / BC levels              */
    int64_t rows_per_acc[GK_NUM_ACC];     /* tasks per accumulator               */
    double  acc_ms[GK_NUM_KERNELS];       /* device time per accumulator kernel  */
    double  acc_bytes[GK_NUM_KERNELS];    /* algorithmic bytes per kernel        */
    int64_t acc_launches[GK_NUM_KERNELS]; /* launches per kernel                 */
    double  bytes_alg;                /* algorithmic bytes, all product kernels  */
    double  flop;                     /* streamed B entries (sum_k |B_k|)         */
    double  hits;                     /* mask hits (triangles for TC)            */
} gk_timing;
int gk_last_timing(gk_ctx *ctx, gk_timing *out);

/* ---- host graphs (construction and cleaning, host threads) ---------------
 * A gk_hgraph holds a cleaned undirected simple graph: symmetric pattern A
 * (rows sorted, no self loops, no duplicates) and its strictly-lower part L
 * (extract_lower_triangle, include/matrix.h:698-726).  With relabel=1 the
 * vertices are renumbered by (degree descending, old id ascending).        */
typedef struct gk_hgraph gk_hgraph;
int  gk_hgraph_rmat(int scale, int edge_factor, double a, double b, double c,
                    uint64_t seed, int relabel, gk_hgraph **out);
int  gk_hgraph_er(int64_t n, double mean_degree, uint64_t seed, int relabel,
                  gk_hgraph **out);
int  gk_hgraph_chung_lu(int64_t n, double exponent, double mean_degree,
                        uint64_t seed, int relabel, gk_hgraph **out);
/* Many graphs as one block-diagonal graph: graph g is Erdos-Renyi (kind 0,
 * mean degree mean_degree[g]) or Chung-Lu (kind 1, exponent[g] and
 * mean_degree[g]) on n[g] vertices with seed[g], built and cleaned like
 * gk_hgraph_er / gk_hgraph_chung_lu (graphs in parallel on host threads),
 * and placed on vertices [voff[g], voff[g+1]) of the union; voff gets
 * ngraphs + 1 entries.                                                      */
int  gk_hgraph_batch(int64_t ngraphs, const int32_t *kind, const int64_t *n,
                     const double *exponent, const double *mean_degree,
                     const uint64_t *seed, int relabel, gk_hgraph **out,
                     int64_t *voff);
int  gk_hgraph_grid(int64_t rows, int64_t cols, int diagonal, int relabel,
                    gk_hgraph **out);
int  gk_hgraph_from_edges(int64_t n, int64_t ne, const int32_t *u,
                          const int32_t *v, int relabel, gk_hgraph **out);
/* Pattern of an arbitrary CSR (e.g. from gk_mtx_read): keeps rows as the
 * reference does -- lower part mirrored (build_symm_from_lower(
 * extract_lower_triangle(A)), src/main.cpp:214-215) but sorted and deduped. */
int  gk_hgraph_from_csr(const gk_csr *A, int relabel, gk_hgraph **out);
int  gk_hgraph_view(const gk_hgraph *g, gk_csr *A, gk_csr *L);
/* new id of old vertex v (relabel permutation); NULL-safe when relabel=0 */
int  gk_hgraph_perm(const gk_hgraph *g, int32_t *new_of_old);
void gk_hgraph_free(gk_hgraph *g);

/* ---- Matrix Market reader (reference semantics) -------------------------
 * include/matrix.h:259-395: 1-based coordinates, self loops dropped,
 * "symmetric" entries mirrored, column order within a row = file order,
 * duplicates kept; complex/array and non-square rejected.  Values come back
 * as double (pattern files give 1.0).                                       */
typedef struct gk_mtx gk_mtx;
int  gk_mtx_read(const char *path, gk_mtx **out);
int  gk_mtx_view(const gk_mtx *mt, gk_csr *pattern, const double **val,
                 char matcode[4]);
void gk_mtx_free(gk_mtx *mt);

/* The reference's other formats (spMtx(filename, format),
 * include/matrix.h:44-61; readers :396-533, writers :209-257), into the same
 * gk_mtx: "mtx" (as gk_mtx_read), "crs" (text: m nz matcode, columns,
 * values, row pointers), "bin" (4 matcode bytes, m n nz as 64-bit sizes,
 * int32 row pointers, columns, values), "graph" (text: m nz, then one line
 * of 0-based columns per row), "rmat" (text: m n nz, then 0-based pairs;
 * both directions stored, rows sorted and deduplicated).  Deviations from
 * the reference: crs sets n = m (the reference leaves it 0), graph reads
 * row i from line i after the header (the reference's first getline takes
 * the rest of the header line as row 0), rmat rows come out sorted and
 * deduplicated.                                                             */
int gk_matrix_read(const char *path, const char *format, gk_mtx **out);
/* Writers "bin" (each array with its own data: the reference's writer puts
 * the row pointers in all three, include/matrix.h:220-222), "mtx" (general
 * coordinate, integer values or pattern when val is NULL) and "crs".  val
 * (NULL: pattern, written as ones in bin/crs) holds int32 values per entry,
 * e.g. edge supports or truss numbers; matcode NULL picks MCIG / MCPG.     */
int gk_matrix_write(const char *path, const char *format, const gk_csr *A,
                    const int32_t *val, const char matcode[4]);

/* ---- device-resident CSR ----------------------------------------------- */
typedef struct gk_dcsr gk_dcsr;
int  gk_dcsr_upload(gk_ctx *ctx, const gk_csr *h, gk_dcsr **out);
void gk_dcsr_free(gk_dcsr *d);
int64_t gk_dcsr_nnz(const gk_dcsr *d);
int  gk_dcsr_dims(const gk_dcsr *d, int64_t *m, int64_t *n, int64_t *nz);
/* Copies a device CSR's row pointers (m+1) and columns (nz) to host arrays. */
int  gk_dcsr_download(gk_ctx *ctx, const gk_dcsr *d, int32_t *rst, int32_t *col);

/* ---- on-device graph construction ----------------------------------------
 * The gk_hgraph_* generators and cleaning pipeline on the GPU: the same
 * graphs bit for bit (counter-based generators, sort-based cleaning), left
 * in HBM as *A (symmetric, sorted rows, no self loops or duplicates) and
 * *L (strict lower part), both freed with gk_dcsr_free.  d_perm (device, n
 * entries, or NULL) receives the new id of every generated vertex.  This is
 * the step before the path: the reference builds its CSR on the host
 * (build_adjacency_matrix / extract_lower_triangle, include/matrix.h:643-726). */
int gk_dgraph_rmat(gk_ctx *ctx, int scale, int edge_factor, double a, double b,
                   double c, uint64_t seed, int relabel, gk_dcsr **A,
                   gk_dcsr **L, int32_t *d_perm);
int gk_dgraph_er(gk_ctx *ctx, int64_t n, double mean_degree, uint64_t seed,
                 int relabel, gk_dcsr **A, gk_dcsr **L, int32_t *d_perm);
int gk_dgraph_chung_lu(gk_ctx *ctx, int64_t n, double exponent,
                       double mean_degree, uint64_t seed, int relabel,
                       gk_dcsr **A, gk_dcsr **L, int32_t *d_perm);
int gk_dgraph_grid(gk_ctx *ctx, int64_t rows, int64_t cols, int diagonal,
                   int relabel, gk_dcsr **A, gk_dcsr **L, int32_t *d_perm);

/* ---- masked sparse-sparse product C<M> = A*B ----------------------------
 * Output pattern := pattern of M (include/matrix_la.h:564-568); c_val[p]
 * for p in [0, M.nz) is sum_k A(i,k)*B(k,j) at M's entry p = (i,j);
 * explicit zeros kept.  Requires A.n == B.m, A.m == M.m, B.n == M.n and
 * sorted, duplicate-free columns in M.  int32 wrap-around like the
 * reference's int arithmetic.                                               */
int gk_masked_spgemm_i32(gk_ctx *ctx, const gk_csr *A, const gk_csr *B,
                         const gk_csr *M, int32_t *c_val, const gk_opts *o);
int gk_masked_spgemm_dev(gk_ctx *ctx, const gk_dcsr *A, const gk_dcsr *B,
                         const gk_dcsr *M, int32_t *d_cval, const gk_opts *o);

/* ---- complement-mask product C<not M> = A*B ------------------------------
 * mxmm_msa_cmask (include/matrix_la.h:733-934): C holds every (i,j) that a
 * product term A(i,k)*B(k,j) reaches and M does not hold (C.m = M.m, C.n =
 * M.n), with the int32 (wrap-around) sum of the terms, explicit zeros kept,
 * rows in column order.  The pattern is not known beforehand, so *C is a
 * library-built device CSR with values (gk_dcsr_dims, gk_dcsr_download,
 * gk_dcsr_download_val; freed with gk_dcsr_free): with *C NULL on entry a
 * new one, with *C a result of an earlier call that one refilled, its
 * storage grown only when too small (as the reference's resizeVals grows
 * C).  M needs sorted rows.  Rows go one of two ways (the result does not depend on it): sorted
 * (product terms expanded, radix-sorted and summed) or dense (a bitmap and
 * sums per window of columns in LDS; needs B's rows sorted when C is wider
 * than one window).  dense_rows: GK_CMASK_AUTO picks per row by its terms
 * against its A entries and C's width, GK_CMASK_SORTED and GK_CMASK_DENSE
 * force one way where allowed.  Dense rows are written in one pass to a
 * work pool of up to min(terms, C.n) entries per row and copied into C
 * when that pool stays under 8 GiB, else counted in a first pass and
 * written in a second; GK_CMASK_DENSE_COUNTED forces every row dense and
 * counted (no pool).  gk_last_timing afterwards: kernel_ms = the
 * term expansion, select_ms = sort and reduce, acc_ms[0] / acc_bytes[0] =
 * the dense rows' two kernels' time and algorithmic bytes (entries written
 * to a pool then copied, or a count pass then the write), launches = row chunks,
 * iterations = dense rows, flop = product terms, hits = entries of C,
 * bytes_alg = the expansion's algorithmic bytes.                          */
enum gk_cmask_rows { GK_CMASK_AUTO = -1, GK_CMASK_SORTED = 0, GK_CMASK_DENSE = 1, GK_CMASK_DENSE_COUNTED = 2 };
int gk_masked_spgemm_cmask_dev(gk_ctx *ctx, const gk_dcsr *A, const gk_dcsr *B,
                               const gk_dcsr *M, int32_t dense_rows, gk_dcsr **C);
int gk_dcsr_download_val(gk_ctx *ctx, const gk_dcsr *d, int32_t *val);

/* ---- triangle counting C<L> = L*L, sum of C -----------------------------
 * src/main.cpp:69-90; L strictly lower, sorted columns.                    */
int gk_tc_count(gk_ctx *ctx, const gk_csr *L, uint64_t *count, const gk_opts *o);
int gk_tc_count_dev(gk_ctx *ctx, const gk_dcsr *L, uint64_t *count, const gk_opts *o);
/* per-row triangle counts (sum over the mask row), length L.m */
int gk_tc_rows_dev(gk_ctx *ctx, const gk_dcsr *L, uint64_t *d_rowcount, const gk_opts *o);

/* ---- triangles through every vertex ---------------------------------------
 * triangle_counting_vertex (src/main.cpp:16-39): counts[i] = (the int sum
 * of row i of SQ<A> = A*A) >> 1, with A's values; A square with sorted rows
 * (it is the mask).  For a symmetric pattern this is the number of
 * triangles through vertex i.  _dev: A resident, counts in device memory
 * (length A.m).                                                           */
int gk_tc_vertex(gk_ctx *ctx, const gk_csr *A, int32_t *counts, const gk_opts *o);
int gk_tc_vertex_dev(gk_ctx *ctx, const gk_dcsr *A, int32_t *d_counts, const gk_opts *o);

/* ---- supports S<A> = A*A at A's pattern (symmetric pattern A) ----------- */
int gk_ktruss_support(gk_ctx *ctx, const gk_csr *A, uint32_t *support, const gk_opts *o);

/* ---- single-k truss, reference semantics (src/main.cpp:93-141) ----------
 * Repeats S<C>=C*C, keeps entries with S >= k-2, until no entry is removed.
 * Output: pattern CSR of the k-truss (rst_out[m+1], col_out[<=A.nz]).      */
int gk_ktruss(gk_ctx *ctx, const gk_csr *A, int k, int32_t *rst_out,
              int32_t *col_out, int64_t *nz_out, int64_t *iterations,
              const gk_opts *o);

/* ---- full truss decomposition -------------------------------------------
 * truss[p] for every entry p of the symmetric pattern A: the largest k such
 * that the edge lies in the k-truss (2 for edges in no triangle).           */
int gk_truss_decompose(gk_ctx *ctx, const gk_csr *A, uint32_t *truss,
                       int32_t *kmax, const gk_opts *o);
int gk_truss_decompose_dev(gk_ctx *ctx, const gk_dcsr *A, uint32_t *d_truss,
                           int32_t *kmax, const gk_opts *o);

/* ---- batched Brandes betweenness, FP64 ----------------------------------
 * include/betweenness_centrality.h:406-506 with an explicit source list:
 * bc[v] = sum over sources s of delta_s(v) (unnormalised, delta_s(s) = 0),
 * sigma and delta kept in FP64.  Sources are processed in batches of up to
 * 256 columns.  A is the (possibly directed) adjacency pattern; forward
 * pulls over A^T, backward over A, as the reference does.                   */
int gk_bc_batch_f64(gk_ctx *ctx, const gk_csr *A, const int32_t *sources,
                    int64_t nsrc, double *bc, const gk_opts *o);
int gk_bc_batch_dev(gk_ctx *ctx, const gk_dcsr *A, const gk_dcsr *AT,
                    const int32_t *sources, int64_t nsrc, double *d_bc,
                    const gk_opts *o);

/* ---- allocation-free BC ---------------------------------------------------
 * No counterpart in GraphKokkos.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc): gk_bc_batch_dev's sweeps and results with every work
 * buffer in a fixed array of GK_BC_NOALLOC_SCRATCH_BYTES per GPU, which the
 * library holds from gk_init until the process ends; the call itself
 * allocates nothing.  The caller passes A (and AT), the sources and d_bc.
 * Limits: A's rows sorted without repeats (as gk_dcsr_upload records);
 * AT given, or A symmetric (no transpose is built); n vertices need
 * n x (2 KiB + 40 B) plus n x 32 B per BFS level, levels 0 to the first
 * empty one (about 55,000 vertices at 8 levels), GK_ERANGE beyond that.  Calls on one GPU share its array, so
 * calls that overlap on one GPU race.  gk_last_timing reports the levels
 * only (no event timers).                                                  */
#define GK_BC_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
int gk_bc_batch_dev_noalloc(gk_ctx *ctx, const gk_dcsr *A, const gk_dcsr *AT,
                            const int32_t *sources, int64_t nsrc, double *d_bc);

/* ---- batching: many independent graphs in one launch sequence -----------
 * A is the block-diagonal union of ngraphs simple undirected graphs (graph g
 * on vertices [voff[g], voff[g+1]), voff[0] = 0, voff[ngraphs] = A.m; an
 * edge joining two graphs is rejected).  One step computes, for every
 * graph, its triangle count (src/main.cpp:69-90) and its k-truss
 * (src/main.cpp:93-141): one support product over the union, a per-graph
 * reduction, and one persistent kernel that peels the graphs largest first,
 * one workgroup per graph.  use_graph = 1 captures the step once as a
 * hipGraph and replays it `reps` times; 0 launches it eagerly each time.
 * Outputs (after the last repetition): triangles[g], truss_edges[g] (edges
 * left in graph g's k-truss) and, when d_keep is not NULL, d_keep[p] = 1 for
 * the entries p of A in the k-truss.  ms_per_rep (optional) gets the mean
 * device time of one step; rep_digest (optional, reps entries) a hash of
 * every repetition's per-graph results.  No counterpart in GraphKokkos,
 * which handles one graph per run.                                          */
int gk_batch_tc_ktruss_dev(gk_ctx *ctx, const gk_dcsr *A, int64_t ngraphs,
                           const int64_t *voff, int k, int use_graph, int reps,
                           uint64_t *triangles, int64_t *truss_edges,
                           uint8_t *d_keep, double *ms_per_rep,
                           uint64_t *rep_digest, const gk_opts *o);

/* ---- learned accumulator selector ----------------------------------------
 * A 16 -> 64 -> 64 -> 4 FP32 MLP predicting each accumulator's cost (log2
 * ns of device time) of a row from 16 row features; GK_SEL_LEARNED picks the
 * cheapest prediction per row.  No counterpart in GraphKokkos (its products
 * take the accumulator from the caller: mxmm_msa vs mxmm_mca).
 * Parameter vector: W1[16x64] b1[64] W2[64x64] b2[64] W3[64x4] b3[4]
 * feature mean[16] feature std[16] (gk_model_num_params() floats).        */
int gk_model_num_params(void);
int gk_model_set(gk_ctx *ctx, const float *params, int n);
int gk_model_get(gk_ctx *ctx, float *params, int n);
int gk_model_clear(gk_ctx *ctx);
/* Y[n x 4] = model(X[n x 16]) */
int gk_model_forward(gk_ctx *ctx, const float *X, int64_t n, float *Y);
/* loss = mean over valid outputs of (Y - T)^2; grad of W1..b3 (5508 floats) */
int gk_model_grad(gk_ctx *ctx, const float *X, const float *T, const uint8_t *valid,
                  int64_t n, float *grad, float *loss);
/* Adam on mini-batches from a seeded initialisation; loads the result. */
int gk_model_train(gk_ctx *ctx, const float *X, const float *T, const uint8_t *valid,
                   int64_t n, int epochs, int batch, float lr, uint64_t seed,
                   float *final_loss);
/* Training samples of C<M> = A*B (mode 0: triangle counting, 1: values):
 * X[m x 16] features, T[m x 4] measured log2 cost per accumulator, valid[m x 4].
 * The rows are timed under opts (NULL: defaults), so the engine knobs the
 * selector will run with (hub_window, thresholds) shape the labels.
 * opts.selector_labels = 1: MSA-LDS and MCA-wave run concurrently in two
 * products, each row in one of them by turns, and a label is the row's
 * wave-time (clock64 cycles of its tasks times the waves a task holds, in ns
 * at 2.4 GHz); hash-global and dense-global are forced in turn, same units.
 * 0: every accumulator forced in turn, a label being the row's share of its
 * kernel's measured time. */
int gk_selector_samples(gk_ctx *ctx, const gk_dcsr *A, const gk_dcsr *B,
                        const gk_dcsr *M, int mode, float *X, float *T,
                        uint8_t *valid, const gk_opts *opts);

/* ---- device buffers for the *_dev entry points ------------------------- */
int  gk_dmalloc(gk_ctx *ctx, size_t bytes, void **dptr);
void gk_dfree(gk_ctx *ctx, void *dptr);
int  gk_memcpy_d2h(gk_ctx *ctx, void *dst, const void *src, size_t bytes);
int  gk_memcpy_h2d(gk_ctx *ctx, void *dst, const void *src, size_t bytes);
int  gk_memcpy_d2d(gk_ctx *ctx, void *dst, const void *src, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif /* GRAPHHIP_H */
