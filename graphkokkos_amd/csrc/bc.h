// Host side of the batched Brandes sweeps (bc.hip), shared by the entry
// points that run them: the work buffers of one call come from a Pool.
#pragma once

#include <vector>

#include "gk_device.h"

namespace gk {
namespace bc {

// sources per batch (one bit / one FP64 column each)
constexpr int kSrc = 256;

// Device buffers of one call, handed out in the order bc_sweeps asks for
// them.  level(d) is the frontier bitmap of BFS level d (n x 4 words),
// made on first use.
struct Pool {
    virtual ~Pool() = default;
    virtual int get(size_t bytes, void **p) = 0;
    virtual int level(int d, uint64_t **p) = 0;
    template <class T>
    int alloc(T **p, size_t n) { return get(sizeof(T) * (n ? n : 1), (void **)p); }
};

// Forward and backward sweeps of every batch of kSrc sources on A (square,
// device CSR), with trst/tcol its transpose (A itself for a symmetric
// pattern); d_bc (n entries) receives the scores.  dups holds the repeated
// neighbours of each batch's source rows as (column in the batch,
// neighbour, multiplicity) triples, or is null when A's rows are sorted.
// Event timers of the forward and backward parts go to ctx->last when
// `timed`.
int sweeps(gk_ctx *ctx, const gk_dcsr *A, const int32_t *trst, const int32_t *tcol, const int32_t *sources,
           int64_t nsrc, double *d_bc, Pool &P, const std::vector<std::vector<int32_t>> *dups, bool timed);

// Argument checks shared by the entry points.
int check_args(gk_ctx *ctx, const gk_dcsr *A, const gk_dcsr *AT, const int32_t *sources, int64_t nsrc,
               const double *d_bc);

}  // namespace bc
}  // namespace gk
