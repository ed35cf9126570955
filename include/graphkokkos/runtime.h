// graphkokkos/runtime.h -- process-wide device context behind the C++
// drop-in headers (graphkokkos/matrix.h, matrix_la.h, algorithms.h,
// betweenness_centrality.h).
//
// GraphKokkos needs Kokkos::initialize before any product runs
// (src/main.cpp:358).  Here gk::initialize binds one GPU and creates the
// libgraphhip context every drop-in call uses; gk::ScopeGuard mirrors
// Kokkos::ScopeGuard.  Calls made before gk::initialize bind device 0 on
// first use.  Failures of the library throw gk::error.
#pragma once

#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>

#include "../graphhip.h"

namespace gk {

class error : public std::runtime_error {
public:
    error(int rc, const std::string &what) : std::runtime_error(what), code(rc) {}
    int code;
};

inline void check(int rc, const char *where) {
    if (rc != GK_OK) throw error(rc, std::string(where) + " failed: " + gk_last_error());
}

namespace detail {
struct Runtime {
    std::mutex mu;
    gk_ctx *ctx = nullptr;
    int device = 0;
    gk_opts opts{};
    bool opts_set = false;
};
inline Runtime &runtime() {
    static Runtime r;
    return r;
}
}  // namespace detail

inline void initialize(int device = 0) {
    auto &r = detail::runtime();
    std::lock_guard<std::mutex> g(r.mu);
    if (r.ctx) return;
    check(gk_init(device, &r.ctx), "gk_init");
    r.device = device;
}

inline void finalize() {
    auto &r = detail::runtime();
    std::lock_guard<std::mutex> g(r.mu);
    if (r.ctx) gk_finalize(r.ctx);
    r.ctx = nullptr;
}

inline bool is_initialized() { return detail::runtime().ctx != nullptr; }

// The context the drop-in calls run on (device 0 unless initialize chose one).
inline gk_ctx *context() {
    if (!is_initialized()) initialize(0);
    return detail::runtime().ctx;
}

// Options handed to every drop-in product: accumulator selector, forced
// accumulator and thresholds (gk_opts in graphhip.h).  Defaults otherwise.
inline const gk_opts &options() {
    auto &r = detail::runtime();
    if (!r.opts_set) {
        gk_opts_default(&r.opts);
        r.opts_set = true;
    }
    return r.opts;
}
inline void set_options(const gk_opts &o) {
    auto &r = detail::runtime();
    r.opts = o;
    r.opts_set = true;
}

struct ScopeGuard {
    explicit ScopeGuard(int device = 0) { initialize(device); }
    ~ScopeGuard() { finalize(); }
    ScopeGuard(const ScopeGuard &) = delete;
    ScopeGuard &operator=(const ScopeGuard &) = delete;
};

}  // namespace gk
