// Shared declarations of the masked-product engine (spgemm.hip) used by the
// k-truss, selector and batching translation units.
#pragma once

#include "gk_device.h"

namespace gk {

// Product modes.
enum { MODE_TC = 0, MODE_VALUES = 1, MODE_SUPPORT = 2 };

// Capacities (entries).  A task covers at most kMaxTaskA A-entries so its
// B-row prefix fits LDS; MSA-LDS holds mask rows up to kLdsMaxMask in a
// hash of kLdsHashCap slots; MCA-wave holds mask rows up to kMcaMaxMask.
constexpr int kMaxTaskA = 256;
// MCA-wave tasks hold at most kMcaMaxTaskA A entries (four tasks share a
// wave's LDS).
constexpr int kMcaMaxTaskA = 64;
constexpr int kLdsHashCap = 4096;
constexpr int kLdsMaxMask = 1536;
constexpr int kMcaMaxMask = 64;
constexpr int kHashGlobalMaxMask = 1 << 20;
// MSA-LDS runs as two kernels: mask rows up to kLdsSmallMask in a table of
// kLdsSmallCap slots (small LDS footprint, high occupancy) and longer rows in
// the kLdsHashCap table.  The long-row kernel is an internal class of its own.
constexpr int kLdsSmallMask = 128;
constexpr int kLdsSmallCap = 512;
// Triangle counting: MSA-LDS keeps mask columns below kWindow in an LDS
// bitmap and the rest in a hash of kLdsSmallCap / kWindowHashLong slots.
constexpr int kWindow = 32768;
constexpr int kWindowHashLong = 2048;
// long mask rows: one workgroup of kWindowWavesLong waves per task
constexpr int kWindowWavesLong = 4;
constexpr int kWindowBlocksLong = 8;  // per CU
// streamed entries per lane in flight in the window kernels
constexpr int kWindowU = 8;
// ... in the MCA-wave kernel
constexpr int kMcaU = 8;
// rounds (of 64 / 16 tasks) an MCA-wave wave takes per ticket, at least
constexpr int kMcaRounds = 8;
// Task tickets: a class with many tasks per workgroup takes several
// consecutive tasks per ticket, so that each workgroup (MCA-wave: each
// wave) draws about kTicketsPerUnit tickets: every ticket is an atomic on
// one counter that the class's whole grid shares, and with millions of
// small tasks those atomics, not the tasks, set the pace.
constexpr int kTicketsPerUnit = 16;
constexpr int kTicketChunkMax = 64;
__host__ __device__ constexpr int32_t ticket_chunk(int32_t ntask, int32_t units, int32_t least) {
    const int32_t k = ntask / (units * kTicketsPerUnit);
    return k < least ? least : (k > kTicketChunkMax ? kTicketChunkMax : k);
}
// triangle counting: MCA-wave's per-task bit filter of the mask row
// (2^kMcaFilterLg bits; a mask row of 64 columns sets at most 1/32 of them)
constexpr int kMcaFilterLg = 11;
constexpr int kMcaFilterWords = (1 << kMcaFilterLg) / 32;
constexpr int kNumClass = GK_KERNEL_MSA_LONG + 1;
constexpr int kClassMsaLarge = GK_NUM_ACC;
constexpr int kFlopBins = 32;
constexpr int kNumBins = kNumClass * kFlopBins;
// Persistent-grid sizes (workgroups of one wave per CU).
constexpr int kLdsSmallBlocksPerCu = 24;
constexpr int kLdsBlocksPerCu = 8;
constexpr int kMcaBlocksPerCu = 8;
constexpr int kGlobalBlocksPerCu = 4;

struct RowStats {
    int64_t *flop;   // words the window kernels read per A row: sum of |B_k|, with
                     // hub rows counted by their bitmap words where that is fewer
    int64_t *sflop;  // sum over A row of |B_k| (what the global kernels stream)
    int64_t *mflop;  // MCA-wave work: hub rows cost the mask entries below k
                     // (one bit test each) where that is fewer than |B_k|
    int32_t *mlen;   // mask row length
    int32_t *alen;   // A row length
    int32_t *bmax;   // max |B_k|
    int32_t *ntask;  // tasks of the row
    int8_t *acc;     // accumulator wanted (learned selector) / kernel class chosen (-1: no work)
};

struct SelectParams {
    int selector, forced;
    int32_t mca_max_mask, lds_max_mask, dense_min_mask;
    int32_t split_flop;
    int32_t small_max_mask;
    int32_t hub;     // hub bitmaps in use (window classes read rows by the cheaper form)
    float margin;    // learned selector: log2 cost gain needed to leave the heuristic
};

struct Task {
    int32_t row, a0, a1, split;
};

// The fixed-threshold rule: accumulator of a mask row of mlen columns.
__device__ __forceinline__ int heuristic_accum(const SelectParams &sp, int32_t mlen) {
    if (mlen <= sp.mca_max_mask) return GK_ACC_MCA_WAVE;
    if (mlen <= sp.lds_max_mask) return GK_ACC_MSA_LDS;
    if (mlen >= sp.dense_min_mask) return GK_ACC_DENSE_GLOBAL;
    return GK_ACC_HASH_GLOBAL;
}

struct Plan {
    int32_t ntasks = 0;
    int32_t acc_tasks[kNumClass] = {};      // per kernel class
    int32_t acc_maxmask[kNumClass] = {};
    Task *tasks = nullptr;
    int32_t *bin_off = nullptr;             // device, kNumBins + 1
    unsigned long long *stats = nullptr;    // device, kNumClass * ST_NUM
    unsigned long long *hits = nullptr;     // device, kNumClass
    int32_t *ticket = nullptr;              // device, kNumClass
};

// Per-accumulator work counters gathered while binning (device, u64):
// [acc][0] tasks, [acc][1] sum of mask lengths, [acc][2] sum of A-row
// lengths, [acc][3] streamed B entries; then hits per accumulator.
enum { ST_TASKS = 0, ST_MASK, ST_AROW, ST_FLOP, ST_NUM };

struct Prob {
    const int32_t *__restrict__ arst, *__restrict__ acol, *__restrict__ aval;
    const int32_t *__restrict__ brst, *__restrict__ bcol, *__restrict__ bval;
    const int32_t *__restrict__ mrst, *__restrict__ mcol;
    int32_t *__restrict__ cval;                 // VALUES: C values; SUPPORT: uint32 supports
    unsigned long long *__restrict__ count;     // TC: total
    unsigned long long *__restrict__ rowcount;  // TC: optional per-row counts
    unsigned long long *__restrict__ hits;      // hits per accumulator
    long long *__restrict__ task_cycles;        // optional: clock64 duration of every task
    const Task *__restrict__ tasks;
    const int32_t *__restrict__ bin_off;
    // per kernel class: tasks taken so far (the window and MCA-wave kernels
    // take tasks by atomic ticket; zeroed before every launch of the classes)
    int32_t *ticket;
    // Triangle counting with hub bitmaps: row k < hub_n of B as hub_w
    // words at hub + k * hub_w (bit c set iff column c is in the row).
    const uint32_t *__restrict__ hub;
    int32_t hub_n, hub_w;
    // A's rows are M's rows (C<L> = L*L): the mask columns below an A entry
    // are the row's entries before it, so a window task may test them as
    // single bits of the entry's hub row (hub_choice) where `probe` times
    // their number is below the row's words and entries; 0 = never
    int32_t probe;
};

// Whether A entry k of a window task reads B row k as bitmap words (the
// bitmaps of all nonempty rows below hub_n exist; MCA-wave tests single
// bits of them).
__host__ __device__ __forceinline__ bool hub_row(int32_t hub_n, int32_t k, int32_t len) {
    return k < hub_n && ((k + 31) >> 5) < len;
}

// How a window task reads B row k of len entries, `below` mask columns
// lying below k: streamed entries (unchanged), or the row's bitmap words
// (hlen words at word offset hbs; hub_row), or, where the mask has fewer
// columns below k than both (weighted by P.probe), single-bit tests of
// those columns in the row's bitmap (plen tests of the row at word offset
// pbs).  A moved row gets len = 0.
__device__ __forceinline__ void hub_choice(const Prob &P, int32_t k, int32_t below, int32_t &len, int32_t &hlen,
                                           int32_t &hbs, int32_t &plen, int32_t &pbs) {
    if (k >= P.hub_n || len == 0) return;
    const int32_t words = (k + 31) >> 5;
    const int32_t cost = P.probe * below;  // a test is a gathered word, a word or entry a streamed one
    const int32_t off = k * P.hub_w;
    // Value selects, not branches that store to different outputs: the
    // branch form left the compiler choosing the output's address at run
    // time, which put both outputs in private (scratch) memory.
    const bool tests = P.probe && cost < len && cost < words;
    const bool bitmap = !tests && words < len;
    plen = tests ? below : plen;
    pbs = tests ? off : pbs;
    hlen = bitmap ? words : hlen;
    hbs = bitmap ? off : hbs;
    len = (tests || bitmap) ? 0 : len;
}

// Launch shape of one product's class kernels and the scratch of the
// global accumulators (setup_classes), so launch_classes only launches.
struct ClassLaunch {
    int grid[kNumClass] = {};
    int grid_lds_large = 0;
    int64_t bn = 0;
    int32_t *hkey = nullptr, *hpos = nullptr, *hacc = nullptr;
    int hlg = 6;
    int32_t *dmap = nullptr, *dacc = nullptr;
    int32_t dmm = 0;
};

struct ProductInfo {
    double total_ms = 0, select_ms = 0, kernel_ms = 0;
    int32_t ntasks = 0;
    int32_t acc_tasks[GK_NUM_ACC] = {};   // per accumulator
    double acc_ms[kNumClass] = {};         // per kernel class
    double acc_bytes[kNumClass] = {};
    int32_t class_tasks[kNumClass] = {};
    double flop = 0, hits = 0;
};

// Adds one product's figures to the context's gk_timing (k-truss rounds
// and repeated products accumulate); reset=true starts a new record.
void record_timing(gk_ctx *ctx, const ProductInfo &pi, bool reset);

// hub_n > 0: rows of B below hub_n are read as bitmaps by the window kernels
// (RowStats.flop counts their words).
// hub_n: hub bitmap rows (0: none); probe: the bit-test weight of hub_choice
int row_stats(gk_ctx *ctx, const gk_dcsr *A, const gk_dcsr *B, const gk_dcsr *M, RowStats *st, int32_t hub_n = 0,
              int32_t probe = 0);
// Bitmaps of rows [0, hub_n) of B (B strictly lower: every column below its
// row) into the context's hub scratch; fills P's hub fields.
int build_hub(gk_ctx *ctx, const gk_dcsr *B, int32_t hub_n, Prob *P);
// Hub rows for C<M> = A*B in the given mode under the options (0: none).
int32_t hub_rows(const gk_dcsr *B, int mode, const gk_opts &o);
int plan_tasks(gk_ctx *ctx, const gk_dcsr *A, RowStats &st, const SelectParams &sp, Plan *plan);
// The options a product of the given mode runs with: the fields left at -1
// (mca_max_mask, hub_probe) or 0 (blocks_long, blocks_mca) take the mode's
// default.  Triangle counting: MCA-wave up to mask length 48, hub probes at
// weight 2, 16 long-row and 16 MCA workgroups per CU (the set autotuned on
// R-MAT scale 22, also the fastest on the 2048^2 grid and the 4M-vertex
// Chung-Lu graph); support and value products: 16, 2, 8, 8.
gk_opts mode_opts(const gk_opts &o, int mode);
SelectParams select_params(const gk_opts &o);
int zero_u32(gk_ctx *ctx, void *p, int64_t n);

// C<M> = A*B in the given mode.  d_out: VALUES -> int32 C values (M.nz),
// SUPPORT -> uint32 supports (M.nz), TC -> unused.  d_count / d_rowcount:
// TC totals (zeroed by the caller).
int masked_product(gk_ctx *ctx, const gk_dcsr *A, const gk_dcsr *B, const gk_dcsr *M, int mode, int32_t *d_out,
                   unsigned long long *d_count, unsigned long long *d_rowcount, const gk_opts &o,
                   ProductInfo *info, long long *task_cycles = nullptr, Plan *plan_out = nullptr);

int setup_classes(gk_ctx *ctx, const Plan &plan, int64_t bn, const gk_opts &o, ClassLaunch *L);
// Plan of C<M> = A*B writing d_out, for products launched repeatedly with
// launch_classes (the plan and the scratch live in the context's scratch
// slots until the next product on the context).
int prepare_product(gk_ctx *ctx, const gk_dcsr *A, const gk_dcsr *B, const gk_dcsr *M, int32_t *d_out,
                    const gk_opts &o, Prob *P, ClassLaunch *L);
// Kernel launches only (capturable in a hipGraph); timers optional.
template <int MODE>
int launch_classes(gk_ctx *ctx, const Prob &P, const ClassLaunch &L, Timer *tm);

// Learned selector (selector.hip): writes st.acc for every row.
int model_select(gk_ctx *ctx, RowStats &st, int64_t m, int same_am, const SelectParams &sp);
// Selector value (internal, outside enum gk_selector) of the concurrent
// label products: row i runs in MCA-wave or MSA-LDS by the parity of a hash
// of i plus `round` (gk_opts.forced_accum); mixed_assign writes st.acc.
constexpr int kSelMixed = 16;
int mixed_assign(gk_ctx *ctx, RowStats &st, int64_t m, int round);

}  // namespace gk
