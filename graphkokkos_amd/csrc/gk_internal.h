// Internal helpers shared by the host and device translation units of
// libgraphhip.so: error reporting, host thread pool, counter-based RNG.
#pragma once

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstdarg>
#include <string>
#include <thread>
#include <vector>
#include <algorithm>

#include "graphhip.h"

namespace gk {

// Thread-local message behind gk_last_error().
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
void clear_error();

// Number of host threads used by construction loops.
int host_threads();

// Chung-Lu inverse-CDF table: vertex i has weight (i+1)^(-1/(exponent-1));
// cdf[i] = normalised prefix sum, cdf[n-1] = 1.  An endpoint drawn from u in
// [0,1) is the first i with cdf[i] > u.
void chung_lu_cdf(int64_t n, double exponent, std::vector<double> &cdf);

// Static block partition of [0, n) over host threads; f(begin, end, tid).
template <class F>
void par_blocks(int64_t n, F &&f) {
    int nt = host_threads();
    if (n < 4096 || nt <= 1) { f(int64_t(0), n, 0); return; }
    nt = (int)std::min<int64_t>(nt, (n + 4095) / 4096);
    std::vector<std::thread> th;
    th.reserve(nt);
    for (int t = 0; t < nt; ++t) {
        int64_t b = n * t / nt, e = n * (t + 1) / nt;
        th.emplace_back([&f, b, e, t] { f(b, e, t); });
    }
    for (auto &x : th) x.join();
}

// Dynamic chunks of [0, n); f(i) per index.  For skewed per-row work.
template <class F>
void par_dynamic(int64_t n, int64_t chunk, F &&f) {
    int nt = host_threads();
    if (n <= chunk || nt <= 1) { for (int64_t i = 0; i < n; ++i) f(i); return; }
    std::atomic<int64_t> next{0};
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&] {
            for (;;) {
                int64_t b = next.fetch_add(chunk);
                if (b >= n) break;
                int64_t e = std::min(n, b + chunk);
                for (int64_t i = b; i < e; ++i) f(i);
            }
        });
    for (auto &x : th) x.join();
}

// SplitMix64 finaliser; rng(seed, stream, counter) is a counter-based
// generator, so edge e of a graph draws the same numbers on any thread.
static inline uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
static inline uint64_t rng(uint64_t seed, uint64_t stream, uint64_t ctr) {
    return mix64(mix64(seed ^ (stream * 0xd1b54a32d192ed03ull)) + ctr * 0x9e3779b97f4a7c15ull);
}
static inline double u01(uint64_t x) { return double(x >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace gk

#define GK_CHECK_ARG(cond, ...)                 \
    do {                                        \
        if (!(cond)) {                          \
            gk::set_error(__VA_ARGS__);         \
            return GK_EINVAL;                   \
        }                                       \
    } while (0)
