// Learned accumulator selector: a 16 -> 64 -> 64 -> 4 FP32 MLP that
// predicts, for every row of a masked product, the cost of each accumulator
// (MSA-LDS, MCA-wave, hash-global, dense-global) in log2 ns of device time; the row
// takes the accumulator with the smallest prediction.  Rows the chosen
// accumulator cannot hold fall back by the same capacity rules as the
// fixed-threshold heuristic, which stays the default selector.
//
// Training data come from the product itself: for every accumulator the
// product runs with that accumulator forced and every task records its
// clock64 duration; a row's label (log2 ns) is its share of its class
// kernel's measured time, split in proportion to the rows' task durations.
//
// Kernels, all hand-written:
//   k_features      16 features per row from the row statistics
//   k_mlp_select    features -> MLP forward -> argmin, weights in LDS,
//                   hidden layers on FP32 MFMA, 32 rows per wave
//   k_mlp_forward   the same forward over a feature matrix (evaluation)
//   k_mlp_grad      masked mean-squared-error loss and its gradient,
//                   per-block LDS accumulation then one atomic per weight
//   k_adam          Adam update
//   k_row_cycles    per-row sums of task durations (labels: wave-time of
//                   the row, or its share of the class kernel's time)
//   k_mixed_assign  MSA-LDS / MCA-wave by turns for the concurrent labels
#include <cmath>
#include <cstring>
#include <vector>

#include "gk_device.h"
#include "spgemm.h"

using namespace gk;

namespace gk {

constexpr int kIn = 16, kH = 64, kOut = GK_NUM_ACC;
// parameter layout (floats)
constexpr int oW1 = 0, oB1 = oW1 + kIn * kH, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH,
              oB3 = oW3 + kH * kOut, oMean = oB3 + kOut, oStd = oMean + kIn, kParams = oStd + kIn;
constexpr int kTrainable = oMean;  // mean/std are fixed from the training set

struct Model {
    float *params = nullptr;  // device, kParams
    bool ready = false;
};

void model_free(Model *m) {
    if (!m) return;
    if (m->params) (void)hipFree(m->params);
    delete m;
}

static int model_get(gk_ctx *ctx, Model **out) {
    if (!ctx->model) {
        ctx->model = new Model();
        if (hipMalloc(&ctx->model->params, sizeof(float) * kParams) != hipSuccess) {
            set_error("hipMalloc for the selector model failed");
            return GK_ENOMEM;
        }
    }
    *out = ctx->model;
    return GK_OK;
}

// ---- features ----------------------------------------------------------------
// log2-scaled sizes plus a one-hot of the mask length's range:
//  0 log2(1+mask)   1 log2(1+|A row|)   2 log2(1+flop)   3 log2(1+max |B_k|)
//  4 log2(1+flop/|A row|)  5 log2(1+flop/mask)  6 log2(1+mask)-log2(1+|A row|)
//  7 1 if A's pattern is the mask's (C<L> = L*L style products)
//  8..15 mask length in (0,8], (8,32], (32,64], (64,128], (128,512],
//        (512,1536], (1536,16384], above
__device__ __forceinline__ void row_features(const RowStats &st, int64_t i, int same_am, float *f) {
    const float ml = (float)st.mlen[i], al = (float)st.alen[i], fl = (float)st.flop[i], bm = (float)st.bmax[i];
    f[0] = log2f(1.f + ml);
    f[1] = log2f(1.f + al);
    f[2] = log2f(1.f + fl);
    f[3] = log2f(1.f + bm);
    f[4] = log2f(1.f + fl / fmaxf(al, 1.f));
    f[5] = log2f(1.f + fl / fmaxf(ml, 1.f));
    f[6] = f[0] - f[1];
    f[7] = (float)same_am;
    const int32_t m = st.mlen[i];
    const int32_t lim[7] = {8, 32, 64, 128, 512, 1536, 16384};
    int b = 7;
    for (int x = 0; x < 7; ++x)
        if (m <= lim[x]) { b = x; break; }
#pragma unroll
    for (int x = 0; x < 8; ++x) f[8 + x] = x == b ? 1.f : 0.f;
}

__global__ void k_features(int64_t m, RowStats st, int same_am, float *__restrict__ X) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        float f[kIn];
        row_features(st, i, same_am, f);
#pragma unroll
        for (int x = 0; x < kIn; ++x) X[i * kIn + x] = f[x];
    }
}

// Forward pass of one sample with the weights in LDS; h1/h2 keep the
// post-activation hidden layers for the backward pass.
__device__ __forceinline__ void mlp_forward(const float *__restrict__ w, const float *xin, float *h1, float *h2,
                                            float *y) {
    float x[kIn];
#pragma unroll
    for (int a = 0; a < kIn; ++a) x[a] = (xin[a] - w[oMean + a]) / w[oStd + a];
    for (int j = 0; j < kH; ++j) {
        float s = w[oB1 + j];
#pragma unroll
        for (int a = 0; a < kIn; ++a) s += x[a] * w[oW1 + a * kH + j];
        h1[j] = s > 0.f ? s : 0.f;
    }
    for (int j = 0; j < kH; ++j) {
        float s = w[oB2 + j];
        for (int a = 0; a < kH; ++a) s += h1[a] * w[oW2 + a * kH + j];
        h2[j] = s > 0.f ? s : 0.f;
    }
#pragma unroll
    for (int o = 0; o < kOut; ++o) {
        float s = w[oB3 + o];
        for (int a = 0; a < kH; ++a) s += h2[a] * w[oW3 + a * kOut + o];
        y[o] = s;
    }
}

// ---- inference on matrix cores --------------------------------------------
// The two hidden layers are dense products, [32 x 16] x [16 x 64] and
// [32 x 64] x [64 x 64] per tile of 32 samples, run on FP32 MFMA
// (v_mfma_f32_32x32x2f32: A 32x2, B 2x32, D 32x32 in 16 registers; lane l
// supplies A[l % 32][l / 32] and B[l / 32][l % 32], and holds
// D[8 (v / 4) + 4 (l / 32) + v % 4][l % 32] in register v).  One wave per
// tile; the tile's activations pass through LDS.
typedef float floatx16 __attribute__((ext_vector_type(16)));
constexpr int kTile = 32;
constexpr int kMlpWaves = 4;

struct TileLds {
    float x[kTile][kIn + 1];   // normalised features
    float h[kTile][kH + 1];    // hidden activations (layer 1, then layer 2)
};

__device__ __forceinline__ int d_row(int v, int lane) { return 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3); }

// Orders one wave's LDS writes before its following reads by other lanes
// (a wave's LDS accesses complete in order; this keeps the compiler from
// moving them across).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// y[k] of sample (lane % 32) for k = 0..3 (lanes >= 32 return garbage).
__device__ __forceinline__ void mlp_tile(const float *__restrict__ w, TileLds &t, float *y) {
    const int lane = threadIdx.x & 63, r = lane & 31, hi = lane >> 5;
    floatx16 acc[2] = {floatx16{}, floatx16{}};
    // layer 1: K = 16 in 8 steps of 2, both 32-column halves per step
#pragma unroll 2
    for (int s = 0; s < kIn / 2; ++s) {
        const int k = 2 * s + hi;
        const float a = t.x[r][k];
        acc[0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, w[oW1 + k * kH + r], acc[0], 0, 0, 0);
        acc[1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, w[oW1 + k * kH + 32 + r], acc[1], 0, 0, 0);
    }
    wave_sync();
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const float z = acc[c][v] + w[oB1 + 32 * c + r];
            t.h[d_row(v, lane)][32 * c + r] = z > 0.f ? z : 0.f;
        }
    wave_sync();
    // layer 2: K = 64 in 32 steps of 2
    acc[0] = floatx16{};
    acc[1] = floatx16{};
#pragma unroll 2
    for (int s = 0; s < kH / 2; ++s) {
        const int k = 2 * s + hi;
        const float a = t.h[r][k];
        acc[0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, w[oW2 + k * kH + r], acc[0], 0, 0, 0);
        acc[1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, w[oW2 + k * kH + 32 + r], acc[1], 0, 0, 0);
    }
    wave_sync();
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const float z = acc[c][v] + w[oB2 + 32 * c + r];
            t.h[d_row(v, lane)][32 * c + r] = z > 0.f ? z : 0.f;
        }
    wave_sync();
    // output layer on the vector units: lane r < 32 takes sample r
#pragma unroll
    for (int o = 0; o < kOut; ++o) y[o] = w[oB3 + o];
#pragma unroll 4
    for (int j = 0; j < kH; ++j) {
        const float h = t.h[r][j];
#pragma unroll
        for (int o = 0; o < kOut; ++o) y[o] += h * w[oW3 + j * kOut + o];
    }
    wave_sync();
}

// Rows this small keep the fixed-threshold choice without inference: their
// cost is the per-row overhead whichever accumulator runs them, and on
// graphs of such rows (meshes, the long low-degree tail of power-law
// graphs in degree order) the inference would cost more than any choice
// saves.  A tile of 32 rows all this small skips the network.
constexpr int32_t kTrivialMask = 8;
constexpr int64_t kTrivialFlop = 256;

__global__ __launch_bounds__(64 * kMlpWaves) void k_mlp_select(int64_t m, RowStats st, int same_am,
                                                               SelectParams sp, const float *__restrict__ params) {
    __shared__ float w[kParams];
    __shared__ TileLds tl[kMlpWaves];
    for (int x = threadIdx.x; x < kParams; x += blockDim.x) w[x] = params[x];
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    TileLds &t = tl[wv];
    const int64_t ntiles = (m + kTile - 1) / kTile;
    for (int64_t tile = (int64_t)blockIdx.x * kMlpWaves + wv; tile < ntiles; tile += (int64_t)gridDim.x * kMlpWaves) {
        const int64_t i = tile * kTile + (lane & 31);
        const bool trivial = lane >= kTile || i >= m || (st.mlen[i] <= kTrivialMask && st.flop[i] <= kTrivialFlop);
        if (__ballot(!trivial) == 0) {
            if (lane < kTile && i < m) st.acc[i] = (int8_t)heuristic_accum(sp, st.mlen[i]);
            continue;
        }
        if (lane < kTile) {
            float f[kIn];
            if (i < m) row_features(st, i, same_am, f);
            else
                for (int a = 0; a < kIn; ++a) f[a] = w[oMean + a];
#pragma unroll
            for (int a = 0; a < kIn; ++a) t.x[lane][a] = (f[a] - w[oMean + a]) / w[oStd + a];
        }
        wave_sync();
        float y[kOut];
        mlp_tile(w, t, y);
        if (lane < kTile && i < m) {
            // cheapest predicted accumulator among those that can hold the
            // row; the fixed-threshold choice stays unless it is predicted
            // to cost more than 2^margin times that.  The global
            // accumulators are candidates only for rows too long for
            // MSA-LDS: where a row fits LDS they never win in the measured
            // samples, and their labels (timed alone, HBM-bound) compare
            // badly with those of the LDS kernels timed together.
            const int32_t mlen = st.mlen[i];
            const bool lds = mlen <= kLdsMaxMask;
            const bool can[kOut] = {lds, mlen <= kMcaMaxMask, !lds && mlen <= kHashGlobalMaxMask, !lds};
            int best = lds ? (int)GK_ACC_MSA_LDS : (int)GK_ACC_DENSE_GLOBAL;  // allowed either way
#pragma unroll
            for (int o = 0; o < kOut; ++o)
                if (can[o] && y[o] < y[best]) best = o;
            const int h = heuristic_accum(sp, mlen);
            if (can[h] && y[h] - y[best] <= sp.margin) best = h;
            st.acc[i] = (int8_t)best;
        }
    }
}

__global__ __launch_bounds__(64 * kMlpWaves) void k_mlp_forward(int64_t n, const float *__restrict__ X,
                                                                const float *__restrict__ params,
                                                                float *__restrict__ Y) {
    __shared__ float w[kParams];
    __shared__ TileLds tl[kMlpWaves];
    for (int x = threadIdx.x; x < kParams; x += blockDim.x) w[x] = params[x];
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    TileLds &t = tl[wv];
    const int64_t ntiles = (n + kTile - 1) / kTile;
    for (int64_t tile = (int64_t)blockIdx.x * kMlpWaves + wv; tile < ntiles; tile += (int64_t)gridDim.x * kMlpWaves) {
        const int64_t i = tile * kTile + (lane & 31);
        if (lane < kTile)
#pragma unroll
            for (int a = 0; a < kIn; ++a)
                t.x[lane][a] = i < n ? (X[i * kIn + a] - w[oMean + a]) / w[oStd + a] : 0.f;
        wave_sync();
        float y[kOut];
        mlp_tile(w, t, y);
        if (lane < kTile && i < n)
#pragma unroll
            for (int o = 0; o < kOut; ++o) Y[i * kOut + o] = y[o];
    }
}

// Loss = sum over samples and valid outputs of (y - t)^2 / count; gradient
// of the trainable parameters accumulated into g (zeroed by the caller).
__global__ __launch_bounds__(256) void k_mlp_grad(int64_t n, const float *__restrict__ X, const float *__restrict__ T,
                                                  const uint8_t *__restrict__ V, const int64_t *__restrict__ idx,
                                                  float inv_count, const float *__restrict__ params,
                                                  float *__restrict__ g, float *__restrict__ loss) {
    __shared__ float w[kParams];
    __shared__ float gl[kTrainable];
    for (int x = threadIdx.x; x < kParams; x += blockDim.x) w[x] = params[x];
    for (int x = threadIdx.x; x < kTrainable; x += blockDim.x) gl[x] = 0.f;
    __syncthreads();
    float lsum = 0.f;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = idx ? idx[r] : r;
        float x[kIn], h1[kH], h2[kH], y[kOut], dy[kOut];
        mlp_forward(w, X + i * kIn, h1, h2, y);
#pragma unroll
        for (int a = 0; a < kIn; ++a) x[a] = (X[i * kIn + a] - w[oMean + a]) / w[oStd + a];
#pragma unroll
        for (int o = 0; o < kOut; ++o) {
            const bool v = V[i * kOut + o] != 0;
            const float d = v ? y[o] - T[i * kOut + o] : 0.f;
            lsum += d * d * inv_count;
            dy[o] = 2.f * d * inv_count;
            atomicAdd(&gl[oB3 + o], dy[o]);
        }
        // layer 3 and back into h2
        float dh2[kH];
        for (int a = 0; a < kH; ++a) {
            float s = 0.f;
#pragma unroll
            for (int o = 0; o < kOut; ++o) {
                if (dy[o] != 0.f) atomicAdd(&gl[oW3 + a * kOut + o], h2[a] * dy[o]);
                s += dy[o] * w[oW3 + a * kOut + o];
            }
            dh2[a] = h2[a] > 0.f ? s : 0.f;
        }
        // layer 2 and back into h1
        float dh1[kH];
        for (int a = 0; a < kH; ++a) dh1[a] = 0.f;
        for (int j = 0; j < kH; ++j) {
            if (dh2[j] == 0.f) continue;
            atomicAdd(&gl[oB2 + j], dh2[j]);
            for (int a = 0; a < kH; ++a) {
                if (h1[a] != 0.f) atomicAdd(&gl[oW2 + a * kH + j], h1[a] * dh2[j]);
                dh1[a] += dh2[j] * w[oW2 + a * kH + j];
            }
        }
        // layer 1
        for (int j = 0; j < kH; ++j) {
            const float d = h1[j] > 0.f ? dh1[j] : 0.f;
            if (d == 0.f) continue;
            atomicAdd(&gl[oB1 + j], d);
#pragma unroll
            for (int a = 0; a < kIn; ++a)
                if (x[a] != 0.f) atomicAdd(&gl[oW1 + a * kH + j], x[a] * d);
        }
    }
    __syncthreads();
    for (int x = threadIdx.x; x < kTrainable; x += blockDim.x)
        if (gl[x] != 0.f) atomicAdd(&g[x], gl[x]);
    lsum = dev::wave_sum(lsum);
    if ((threadIdx.x & 63) == 0 && lsum != 0.f) atomicAdd(loss, lsum);
}

__global__ void k_adam(int n, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m1,
                       float *__restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float a = b1 * m1[i] + (1.f - b1) * gi;
        const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
        m1[i] = a;
        m2[i] = b;
        p[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
    }
}

// Label aggregation: clock64 cycles of every row summed over its tasks.
__global__ void k_row_cycles(int32_t lo, int32_t hi, const Task *__restrict__ tasks,
                             const long long *__restrict__ cyc, float per_cu, double *__restrict__ row) {
    for (int32_t t = lo + blockIdx.x * blockDim.x + threadIdx.x; t < hi; t += gridDim.x * blockDim.x)
        atomicAdd(&row[tasks[t].row], (double)cyc[t] / per_cu);
}

// Concurrent label products: row i of round r in MCA-wave or MSA-LDS by the
// parity of a hash of i plus r, so that each row runs in both over two
// rounds and each round runs both kernels on about half of the rows
// (capacity overrides in the binning still apply).
__global__ void k_mixed_assign(int64_t m, RowStats st, int round) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t h = dev::hash32((uint32_t)i ^ (uint32_t)(i >> 32)) >> 31;
        st.acc[i] = (int8_t)(((h + (uint32_t)round) & 1u) ? GK_ACC_MCA_WAVE : GK_ACC_MSA_LDS);
    }
}

// ---- engine hook -----------------------------------------------------------

int mixed_assign(gk_ctx *ctx, RowStats &st, int64_t m, int round) {
    if (m == 0) return GK_OK;
    const int g = (int)std::max<int64_t>(1, std::min<int64_t>((m + 255) / 256, (int64_t)ctx->num_cus * 4));
    k_mixed_assign<<<g, 256, 0, ctx->stream>>>(m, st, round);
    return launch_check("k_mixed_assign");
}

int model_select(gk_ctx *ctx, RowStats &st, int64_t m, int same_am, const SelectParams &sp) {
    if (!ctx->model || !ctx->model->ready) {
        set_error("learned selector requested but no model is loaded (gk_model_set / gk_model_train)");
        return GK_ESTATE;
    }
    if (m == 0) return GK_OK;
    const int g = (int)std::max<int64_t>(1, std::min<int64_t>((m + 255) / 256, (int64_t)ctx->num_cus * 4));
    k_mlp_select<<<g, 256, 0, ctx->stream>>>(m, st, same_am, sp, ctx->model->params);
    return launch_check("k_mlp_select");
}

}  // namespace gk

// ---- C ABI -------------------------------------------------------------------

extern "C" int gk_model_num_params(void) { return gk::kParams; }

extern "C" int gk_model_set(gk_ctx *ctx, const float *params, int n) {
    GK_CHECK_ARG(ctx && params, "null argument");
    GK_CHECK_ARG(n == gk::kParams, "model has %d parameters, got %d", gk::kParams, n);
    for (int a = 0; a < gk::kIn; ++a)
        GK_CHECK_ARG(params[gk::oStd + a] > 0.f, "feature scale %d must be positive", a);
    Model *md;
    GK_TRY(model_get(ctx, &md));
    GK_HIP(hipMemcpyAsync(md->params, params, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    GK_HIP(hipStreamSynchronize(ctx->stream));
    md->ready = true;
    return GK_OK;
}

extern "C" int gk_model_get(gk_ctx *ctx, float *params, int n) {
    GK_CHECK_ARG(ctx && params, "null argument");
    GK_CHECK_ARG(n == gk::kParams, "model has %d parameters, got %d", gk::kParams, n);
    GK_CHECK_ARG(ctx->model && ctx->model->ready, "no model loaded");
    GK_HIP(d2h(ctx, params, ctx->model->params, sizeof(float) * n));
    return GK_OK;
}

extern "C" int gk_model_clear(gk_ctx *ctx) {
    GK_CHECK_ARG(ctx, "null argument");
    if (ctx->model) ctx->model->ready = false;
    return GK_OK;
}

namespace {
struct DevBuf {
    std::vector<void *> p;
    ~DevBuf() { for (void *x : p) (void)hipFree(x); }
    template <class T>
    int alloc(T **out, size_t n) {
        if (hipMalloc((void **)out, sizeof(T) * (n ? n : 1)) != hipSuccess) {
            set_error("hipMalloc(%zu) failed", sizeof(T) * n);
            return GK_ENOMEM;
        }
        p.push_back(*out);
        return GK_OK;
    }
};
}  // namespace

extern "C" int gk_model_forward(gk_ctx *ctx, const float *X, int64_t n, float *Y) {
    GK_CHECK_ARG(ctx && (n == 0 || (X && Y)), "null argument");
    GK_CHECK_ARG(ctx->model && ctx->model->ready, "no model loaded");
    DevBuf b;
    float *dX, *dY;
    GK_TRY(b.alloc(&dX, n * gk::kIn));
    GK_TRY(b.alloc(&dY, n * gk::kOut));
    GK_HIP(hipMemcpyAsync(dX, X, sizeof(float) * n * gk::kIn, hipMemcpyHostToDevice, ctx->stream));
    const int g = (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, (int64_t)ctx->num_cus * 4));
    if (n) k_mlp_forward<<<g, 256, 0, ctx->stream>>>(n, dX, ctx->model->params, dY);
    GK_TRY(launch_check("k_mlp_forward"));
    GK_HIP(d2h(ctx, Y, dY, sizeof(float) * n * gk::kOut));
    return GK_OK;
}

// Loss and gradient of the current model on (X, T, valid): for gradient
// checks against an independent implementation.
extern "C" int gk_model_grad(gk_ctx *ctx, const float *X, const float *T, const uint8_t *valid, int64_t n, float *grad,
                             float *loss) {
    GK_CHECK_ARG(ctx && X && T && valid && grad && loss && n > 0, "null argument");
    GK_CHECK_ARG(ctx->model && ctx->model->ready, "no model loaded");
    DevBuf b;
    float *dX, *dT, *dg, *dl;
    uint8_t *dV;
    GK_TRY(b.alloc(&dX, n * gk::kIn));
    GK_TRY(b.alloc(&dT, n * gk::kOut));
    GK_TRY(b.alloc(&dV, n * gk::kOut));
    GK_TRY(b.alloc(&dg, gk::kTrainable));
    GK_TRY(b.alloc(&dl, 1));
    GK_HIP(hipMemcpyAsync(dX, X, sizeof(float) * n * gk::kIn, hipMemcpyHostToDevice, ctx->stream));
    GK_HIP(hipMemcpyAsync(dT, T, sizeof(float) * n * gk::kOut, hipMemcpyHostToDevice, ctx->stream));
    GK_HIP(hipMemcpyAsync(dV, valid, n * gk::kOut, hipMemcpyHostToDevice, ctx->stream));
    GK_HIP(hipMemsetAsync(dg, 0, sizeof(float) * gk::kTrainable, ctx->stream));
    GK_HIP(hipMemsetAsync(dl, 0, sizeof(float), ctx->stream));
    int64_t cnt = 0;
    for (int64_t i = 0; i < n * gk::kOut; ++i) cnt += valid[i] != 0;
    const float inv = cnt ? 1.f / (float)cnt : 0.f;
    const int g = (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, (int64_t)ctx->num_cus * 2));
    k_mlp_grad<<<g, 256, 0, ctx->stream>>>(n, dX, dT, dV, nullptr, inv, ctx->model->params, dg, dl);
    GK_TRY(launch_check("k_mlp_grad"));
    GK_HIP(d2h(ctx, grad, dg, sizeof(float) * gk::kTrainable));
    GK_HIP(d2h(ctx, loss, dl, sizeof(float)));
    return GK_OK;
}

// Train from scratch (He-normal initialisation from `seed`) with Adam on
// mini-batches; feature mean/std are taken from X.  Returns the last epoch's
// mean loss.
extern "C" int gk_model_train(gk_ctx *ctx, const float *X, const float *T, const uint8_t *valid, int64_t n,
                              int epochs, int batch, float lr, uint64_t seed, float *final_loss) {
    GK_CHECK_ARG(ctx && X && T && valid && n > 0 && epochs > 0 && batch > 0, "bad arguments");
    std::vector<float> p(gk::kParams, 0.f);
    // feature standardisation
    for (int a = 0; a < gk::kIn; ++a) {
        double s = 0, s2 = 0;
        for (int64_t i = 0; i < n; ++i) {
            s += X[i * gk::kIn + a];
            s2 += (double)X[i * gk::kIn + a] * X[i * gk::kIn + a];
        }
        const double mean = s / n, var = std::max(0.0, s2 / n - mean * mean);
        p[gk::oMean + a] = (float)mean;
        p[gk::oStd + a] = var > 1e-12 ? (float)std::sqrt(var) : 1.f;
    }
    // He-normal weights from a counter-based generator; biases 0
    auto normal = [&](uint64_t k) {
        const double u1 = std::max(u01(rng(seed, 7, 2 * k)), 1e-300), u2 = u01(rng(seed, 7, 2 * k + 1));
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    uint64_t ctr = 0;
    for (int i = 0; i < gk::kIn * gk::kH; ++i) p[gk::oW1 + i] = (float)(normal(ctr++) * std::sqrt(2.0 / gk::kIn));
    for (int i = 0; i < gk::kH * gk::kH; ++i) p[gk::oW2 + i] = (float)(normal(ctr++) * std::sqrt(2.0 / gk::kH));
    for (int i = 0; i < gk::kH * gk::kOut; ++i) p[gk::oW3 + i] = (float)(normal(ctr++) * std::sqrt(1.0 / gk::kH));
    // output biases start at the mean label of each output
    for (int o = 0; o < gk::kOut; ++o) {
        double s = 0;
        int64_t c = 0;
        for (int64_t i = 0; i < n; ++i)
            if (valid[i * gk::kOut + o]) { s += T[i * gk::kOut + o]; ++c; }
        p[gk::oB3 + o] = c ? (float)(s / c) : 0.f;
    }
    GK_TRY(gk_model_set(ctx, p.data(), gk::kParams));
    float *params = ctx->model->params;
    DevBuf b;
    float *dX, *dT, *dg, *dl, *m1, *m2;
    uint8_t *dV;
    int64_t *didx;
    GK_TRY(b.alloc(&dX, n * gk::kIn));
    GK_TRY(b.alloc(&dT, n * gk::kOut));
    GK_TRY(b.alloc(&dV, n * gk::kOut));
    GK_TRY(b.alloc(&dg, gk::kTrainable));
    GK_TRY(b.alloc(&dl, 1));
    GK_TRY(b.alloc(&m1, gk::kTrainable));
    GK_TRY(b.alloc(&m2, gk::kTrainable));
    GK_TRY(b.alloc(&didx, n));
    GK_HIP(hipMemcpyAsync(dX, X, sizeof(float) * n * gk::kIn, hipMemcpyHostToDevice, ctx->stream));
    GK_HIP(hipMemcpyAsync(dT, T, sizeof(float) * n * gk::kOut, hipMemcpyHostToDevice, ctx->stream));
    GK_HIP(hipMemcpyAsync(dV, valid, n * gk::kOut, hipMemcpyHostToDevice, ctx->stream));
    GK_HIP(hipMemsetAsync(m1, 0, sizeof(float) * gk::kTrainable, ctx->stream));
    GK_HIP(hipMemsetAsync(m2, 0, sizeof(float) * gk::kTrainable, ctx->stream));
    std::vector<int64_t> perm(n);
    std::vector<float> hl(1);
    float last = 0.f;
    int64_t step = 0;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int ep = 0; ep < epochs; ++ep) {
        // shuffle (Fisher-Yates from the counter-based generator)
        for (int64_t i = 0; i < n; ++i) perm[i] = i;
        for (int64_t i = n - 1; i > 0; --i) {
            const int64_t j = (int64_t)(rng(seed, 11 + ep, i) % (uint64_t)(i + 1));
            std::swap(perm[i], perm[j]);
        }
        GK_HIP(hipMemcpyAsync(didx, perm.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice, ctx->stream));
        GK_HIP(hipMemsetAsync(dl, 0, sizeof(float), ctx->stream));
        for (int64_t b0 = 0; b0 < n; b0 += batch) {
            const int64_t nb = std::min<int64_t>(batch, n - b0);
            int64_t cnt = 0;
            for (int64_t r = b0; r < b0 + nb; ++r)
                for (int o = 0; o < gk::kOut; ++o) cnt += valid[perm[r] * gk::kOut + o] != 0;
            if (!cnt) continue;
            GK_HIP(hipMemsetAsync(dg, 0, sizeof(float) * gk::kTrainable, ctx->stream));
            const int g = (int)std::max<int64_t>(1, std::min<int64_t>((nb + 255) / 256, (int64_t)ctx->num_cus));
            // the loss of the epoch is accumulated with weight nb/n per batch
            k_mlp_grad<<<g, 256, 0, ctx->stream>>>(nb, dX, dT, dV, didx + b0, 1.f / (float)cnt, params, dg, dl);
            ++step;
            const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
            k_adam<<<(gk::kTrainable + 255) / 256, 256, 0, ctx->stream>>>(gk::kTrainable, params, dg, m1, m2, lr, b1,
                                                                         b2, eps, c1, c2);
        }
        GK_TRY(launch_check("model training"));
        GK_HIP(d2h(ctx, hl.data(), dl, sizeof(float)));
        last = hl[0] / (float)((n + batch - 1) / batch);
    }
    if (final_loss) *final_loss = last;
    return GK_OK;
}

// Features (row statistics) and cost labels of every row of C<M> = A*B for
// every accumulator.  X: m x 16; T: m x 4 log2 ns of device time; valid: m x 4
// (0 where the row has no work or the accumulator cannot hold its mask).
extern "C" int gk_selector_samples(gk_ctx *ctx, const gk_dcsr *A, const gk_dcsr *B, const gk_dcsr *M, int mode,
                                   float *X, float *T, uint8_t *valid, const gk_opts *opts) {
    GK_CHECK_ARG(ctx && A && B && M && X && T && valid, "null argument");
    const gk_opts base = mode_opts(resolve_opts(opts), mode == 0 ? MODE_TC : MODE_VALUES);
    GK_CHECK_ARG(mode == 0 || mode == 1, "mode must be 0 (triangle counting) or 1 (values)");
    GK_TRY(require_sorted(M, "M"));
    const int64_t m = A->m;
    const int same_am = A->rst == M->rst && A->col == M->col;
    DevBuf b;
    float *dX;
    double *rowc;
    int8_t *dacc;
    long long *cyc = nullptr;
    GK_TRY(b.alloc(&dX, m * gk::kIn));
    GK_TRY(b.alloc(&rowc, m));
    GK_TRY(b.alloc(&dacc, m));
    // features from the row statistics
    RowStats st;
    GK_TRY(row_stats(ctx, A, B, M, &st, hub_rows(B, mode == 0 ? MODE_TC : MODE_VALUES, base),
                     std::max(base.hub_probe, 0)));
    if (m) {
        const int g = (int)std::max<int64_t>(1, std::min<int64_t>((m + 255) / 256, (int64_t)ctx->num_cus * 4));
        k_features<<<g, 256, 0, ctx->stream>>>(m, st, same_am, dX);
        GK_TRY(launch_check("k_features"));
    }
    GK_HIP(d2h(ctx, X, dX, sizeof(float) * m * gk::kIn));
    int32_t *d_out = nullptr;
    unsigned long long *d_count = nullptr;
    GK_TRY(b.alloc(&d_out, std::max<int64_t>(M->nz, 1)));
    GK_TRY(b.alloc(&d_count, 1));
    std::vector<double> hrow(m);
    std::vector<int8_t> hacc(m);
    if (base.selector_labels) {
        // Concurrent labels: two products in which MSA-LDS and MCA-wave run
        // side by side on alternating halves of the rows (the engine's real
        // schedule: the classes share the CUs), then hash-global and
        // dense-global forced in turn.  A row's label is the wave-time of
        // its tasks: clock64 cycles times the waves a task holds (four for
        // a long-row window task, one for a short-row or global one, a
        // quarter for an MCA-wave tile), in ns at 2.4 GHz, averaged where a
        // row ran in one accumulator more than once.
        std::vector<double> sum((size_t)m * GK_NUM_ACC, 0.0);
        std::vector<int32_t> cnt((size_t)m * GK_NUM_ACC, 0);
        for (int run = 0; run < 4; ++run) {
            gk_opts o = base;
            if (run < 2) {
                o.selector = kSelMixed;
                o.forced_accum = run;
            } else {
                o.selector = GK_SEL_FORCED;
                o.forced_accum = run == 2 ? GK_ACC_HASH_GLOBAL : GK_ACC_DENSE_GLOBAL;
            }
            if (!cyc) GK_TRY(b.alloc(&cyc, A->nz + m + 1));
            GK_HIP(hipMemsetAsync(d_count, 0, 8, ctx->stream));
            Plan plan;
            ProductInfo pi;
            GK_TRY(masked_product(ctx, A, B, M, mode == 0 ? MODE_TC : MODE_VALUES, d_out, d_count, nullptr, o, &pi,
                                  cyc, &plan));
            GK_HIP(d2h(ctx, hacc.data(), st.acc, m));
            GK_HIP(hipMemsetAsync(rowc, 0, sizeof(double) * m, ctx->stream));
            int32_t off[kNumBins + 1];
            GK_HIP(d2h(ctx, off, plan.bin_off, sizeof off));
            for (int c = 0; c < kNumClass; ++c) {
                const int32_t lo = off[c * kFlopBins], hi = off[(c + 1) * kFlopBins];
                // k_row_cycles divides by its `per_cu` argument: 1 / waves per task
                const float waves = c == kClassMsaLarge ? (float)kWindowWavesLong
                                    : c == GK_ACC_MCA_WAVE ? 0.25f : 1.f;
                if (hi > lo)
                    k_row_cycles<<<std::min(1024, (hi - lo + 255) / 256), 256, 0, ctx->stream>>>(
                        lo, hi, plan.tasks, cyc, 1.f / waves, rowc);
            }
            GK_TRY(launch_check("k_row_cycles"));
            GK_HIP(d2h(ctx, hrow.data(), rowc, sizeof(double) * m));
            for (int64_t i = 0; i < m; ++i) {
                const int c = hacc[i];
                if (c < 0 || hrow[i] <= 0) continue;
                const int a = c == kClassMsaLarge ? (int)GK_ACC_MSA_LDS : c;
                if (run >= 2 && a != o.forced_accum) continue;  // a capacity fallback
                sum[(size_t)i * GK_NUM_ACC + a] += hrow[i] / 2.4;
                cnt[(size_t)i * GK_NUM_ACC + a] += 1;
            }
        }
        for (size_t x = 0; x < (size_t)m * GK_NUM_ACC; ++x) {
            valid[x] = cnt[x] > 0;
            T[x] = cnt[x] > 0 ? (float)std::log2(sum[x] / cnt[x]) : 0.f;
        }
        return GK_OK;
    }
    // Forced labels: a row's label is its share of the device time of the
    // class kernel it ran in: the kernel's measured time split over its rows
    // in proportion to their tasks' clock64 durations.  (Durations alone
    // over-count kernels whose tasks wait on each other, such as the four
    // tiles of an MCA wave.)
    for (int a = 0; a < GK_NUM_ACC; ++a) {
        gk_opts o = base;
        o.selector = GK_SEL_FORCED;
        o.forced_accum = a;
        // task durations: one slot per task (tasks <= A entries + rows)
        if (!cyc) GK_TRY(b.alloc(&cyc, A->nz + m + 1));
        GK_HIP(hipMemsetAsync(d_count, 0, 8, ctx->stream));
        Plan plan;
        ProductInfo pi;
        GK_TRY(masked_product(ctx, A, B, M, mode == 0 ? MODE_TC : MODE_VALUES, d_out, d_count, nullptr, o, &pi, cyc,
                              &plan));
        // which class every row really ran in (capacity fallbacks)
        GK_HIP(d2h(ctx, hacc.data(), st.acc, m));
        GK_HIP(hipMemsetAsync(rowc, 0, sizeof(double) * m, ctx->stream));
        int32_t off[kNumBins + 1];
        GK_HIP(d2h(ctx, off, plan.bin_off, sizeof off));
        for (int c = 0; c < kNumClass; ++c) {
            const int32_t lo = off[c * kFlopBins], hi = off[(c + 1) * kFlopBins];
            if (hi > lo)
                k_row_cycles<<<std::min(1024, (hi - lo + 255) / 256), 256, 0, ctx->stream>>>(lo, hi, plan.tasks, cyc,
                                                                                          1.0f, rowc);
        }
        GK_TRY(launch_check("k_row_cycles"));
        GK_HIP(d2h(ctx, hrow.data(), rowc, sizeof(double) * m));
        double cls_cycles[kNumClass] = {};
        for (int64_t i = 0; i < m; ++i)
            if (hacc[i] >= 0) cls_cycles[hacc[i]] += hrow[i];
        double ns_per_cycle[kNumClass] = {};
        for (int c = 0; c < kNumClass; ++c)
            ns_per_cycle[c] = cls_cycles[c] > 0 ? pi.acc_ms[c] * 1e6 / cls_cycles[c] : 0.0;
        for (int64_t i = 0; i < m; ++i) {
            int c = hacc[i];
            const int ran = c == kClassMsaLarge ? (int)GK_ACC_MSA_LDS : c;
            const double ns = c >= 0 ? hrow[i] * ns_per_cycle[c] : 0.0;
            const bool ok = c >= 0 && ran == a && ns > 0;
            valid[i * GK_NUM_ACC + a] = ok;
            T[i * GK_NUM_ACC + a] = ok ? (float)std::log2(ns) : 0.f;
        }
    }
    return GK_OK;
}
