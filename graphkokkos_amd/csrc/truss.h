// Edge-level view of a symmetric pattern, shared by the k-truss
// (ktruss.hip) and batching (batch.hip) translation units.
#pragma once

#include <vector>

#include "gk_device.h"

namespace gk {

// For a sorted symmetric pattern A: the strictly lower part L (edge id =
// position in L), the endpoints of every edge (eu = row > ev = column) and
// the edge id of every entry of A.  Owns its device buffers.
struct EdgeIndex {
    int64_t m = 0, nL = 0;
    int32_t *lrst = nullptr, *lcol = nullptr, *eu = nullptr, *ev = nullptr, *aeid = nullptr;
    std::vector<void *> mem;
    EdgeIndex() = default;
    EdgeIndex(const EdgeIndex &) = delete;
    EdgeIndex &operator=(const EdgeIndex &) = delete;
    ~EdgeIndex() {
        for (void *p : mem) (void)hipFree(p);
    }
    template <class T>
    int alloc(T **p, size_t n) {
        if (hipMalloc((void **)p, sizeof(T) * (n ? n : 1)) != hipSuccess) {
            set_error("hipMalloc(%zu) for an edge index failed", sizeof(T) * n);
            return GK_ENOMEM;
        }
        mem.push_back(*p);
        return GK_OK;
    }
};

int edge_index(gk_ctx *ctx, const gk_dcsr *A, EdgeIndex *X);

int exclusive_scan_i32(gk_ctx *ctx, const int32_t *in, int32_t *out, int64_t n);

}  // namespace gk
