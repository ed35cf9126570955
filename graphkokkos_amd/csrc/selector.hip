// Learned accumulator selector (placeholder until the cost model lands).
#include "gk_device.h"
#include "spgemm.h"

namespace gk {

struct Model {
    int dummy;
};

void model_free(Model *m) { delete m; }

int model_select(gk_ctx *ctx, RowStats &st, int64_t m, int64_t ncols) {
    (void)st; (void)m; (void)ncols;
    if (!ctx->model) {
        set_error("learned selector requested but no model is loaded");
        return GK_ESTATE;
    }
    return GK_OK;
}

}  // namespace gk
