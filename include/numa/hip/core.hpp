#pragma once
// Shared pieces of the numa::hip C++20 layer over the C ABI (include/numa_hip.h):
// error mapping, the device context (and a process-wide default one, so that
// code written against the reference's constructors compiles unchanged),
// device buffers, and the parameter / result structs of the reference's
// solver interface (methods/krylov_newton.hpp:24-49, linalg/gmres.hpp:14-21,
// operators/timing.hpp:16-24).

#include <algorithm>
#include <cstddef>
#include <cstdlib>
#include <limits>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../numa_hip.h"
#include "../operators/operator.hpp"

namespace numa::hip {

using numa::op::DVec;

inline void check(int rc) {
    if (rc == NUMA_OK) return;
    const std::string msg = numa_hip_last_error();
    if (rc == NUMA_ERR_ARG) throw std::invalid_argument(msg);
    if (rc == NUMA_ERR_OVERFLOW) throw std::overflow_error(msg);
    throw std::runtime_error(msg);
}

// One device + stream.  Must outlive every object created on it.
class Context {
public:
    explicit Context(int device = 0) { check(numa_hip_ctx_create(device, &h_)); }
    ~Context() {
        if (h_) numa_hip_ctx_destroy(h_);
    }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    numa_hip_ctx* get() const { return h_; }

private:
    numa_hip_ctx* h_ = nullptr;
};

// Device buffer of n doubles.
class DeviceVec {
public:
    DeviceVec(const Context& c, std::size_t n) : c_(&c), n_(n) {
        check(numa_hip_malloc(c.get(), n * sizeof(double), &p_));
    }
    ~DeviceVec() {
        if (p_) numa_hip_free(c_->get(), p_);
    }
    DeviceVec(const DeviceVec&) = delete;
    DeviceVec& operator=(const DeviceVec&) = delete;
    double* data() const { return static_cast<double*>(p_); }
    std::size_t size() const { return n_; }
    const Context& context() const { return *c_; }
    void upload(const DVec& v) const {
        if (v.size() != n_) throw std::invalid_argument("DeviceVec: size mismatch");
        check(numa_hip_memcpy_h2d(c_->get(), p_, v.data(), n_ * sizeof(double)));
    }
    void download(DVec& v) const {
        v.resize(n_);
        check(numa_hip_memcpy_d2h(c_->get(), v.data(), p_, n_ * sizeof(double)));
    }

private:
    const Context* c_;
    std::size_t n_;
    void* p_ = nullptr;
};

// methods/krylov_newton.hpp:24-38, plus the forcing-mode addition.
struct KrylovParams {
    double rtol_nl = 1e-8;
    std::size_t max_nl = 50;
    double eta_init = 0.5;
    double eta_max = 0.9;
    double gamma = 0.9;
    double alpha = 2.0;
    double eta2_factor = 10.0;
    double eta2_max = 0.5;
    std::size_t gmres_restart = 50;
    std::size_t gmres_max_iters = 500;
    double armijo = 1e-4;
    double min_lambda = 1e-4;
    bool recycle = false;
    int forcing = NUMA_FORCING_EW;  // or NUMA_FORCING_CONSTANT (eta = eta_init)
    // final-solve safeguard (0 = off): eta <= tight_eta once ||F|| <= tight_res ||F(x0)||
    double tight_res = 0.0;
    double tight_eta = 0.0;

    numa_krylov_params c() const {
        return numa_krylov_params{rtol_nl,       (int64_t)max_nl,         eta_init,   eta_max,
                                  gamma,         alpha,                   eta2_factor, eta2_max,
                                  (int64_t)gmres_restart, (int64_t)gmres_max_iters, armijo,
                                  min_lambda,    recycle ? 1 : 0,         forcing,
                                  tight_res,     tight_eta};
    }
};

// methods/krylov_newton.hpp:40-49
struct KrylovResult {
    DVec x;
    bool converged = false;
    std::size_t iters = 0;
    std::size_t n_F = 0, n_JVP = 0, n_D2 = 0;
    std::size_t n_lin_iters = 0;
    std::size_t n_lin_failures = 0;
    std::size_t n_prec = 0;
    double res0 = 0.0, res = 0.0;
    double t_total = 0.0;  // seconds
};

// linalg/gmres.hpp:14-21
struct GmresResult {
    std::size_t iters = 0;
    std::size_t matvecs = 0;
    std::size_t precs = 0;
    double rel_res = 0.0;
    bool converged = false;
    bool breakdown = false;
};

// operators/timing.hpp:16-24: seconds per primitive.  On the device each is
// the GPU time of the primitive's kernels (HIP events on the solver stream).
struct Timers {
    double t_F = 0.0;
    double t_JVP = 0.0;
    double t_D2 = 0.0;
    double t_prec = 0.0;
    double t_setup_dynamic = 0.0;
    void reset() { *this = Timers{}; }
};

// Derivative and precision options of the device solver.
struct SolverOptions {
    int precision = NUMA_PRECISION_FP64;
    bool flexible = false;              // FGMRES
    int jv_mode = NUMA_DERIV_ANALYTIC;  // Bratu3D::jvp as in the reference
    int d2_mode = NUMA_DERIV_ANALYTIC;  // Bratu3D::d2 as in the reference
    // If set, the solve adds its per-primitive times to *timers, as the
    // reference's Timed<Op> / TimedPrecond wrappers do (timing.hpp:37-108).
    Timers* timers = nullptr;
};

// The context used by the constructors that take none (the reference's own
// signatures): device $NUMA_HIP_DEVICE (default 0), created at first use and
// kept until the process ends.
inline const Context& default_context() {
    static std::unique_ptr<Context> ctx;
    static std::once_flag once;
    std::call_once(once, [] {
        const char* d = std::getenv("NUMA_HIP_DEVICE");
        ctx = std::make_unique<Context>(d && *d ? std::atoi(d) : 0);
    });
    return *ctx;
}

}  // namespace numa::hip
