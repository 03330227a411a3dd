#pragma once
// MI355X backend of the matrix-free Bratu3D solver path, as a header-only
// C++20 layer over the C ABI in include/numa_hip.h.
//
// Drop-in for the reference's CPU path (kapumota/chebyshev-nonlinear-lab,
// include/numa): the types below keep the reference's names, members and
// call signatures so that code written against numa::op compiles against
// numa::hip by changing the namespace:
//   numa::hip::Bratu3D          problems/bratu3d.hpp:12-100 (eval / jvp / d2 /
//                               jacobi_diag, satisfies SecondOrderOperator)
//   numa::hip::BratuJacobi      problems/bratu3d.hpp:104-120 (provider)
//   numa::hip::PoissonMGProvider linalg/poisson_mg.hpp:303-320 (provider; the
//                               configuration adds even-n coarsening and the
//                               weighted-Jacobi / Chebyshev smoothers)
//   numa::hip::KrylovParams / KrylovResult   methods/krylov_newton.hpp:24-49
//   numa::hip::GmresResult      linalg/gmres.hpp:14-21
//   numa::hip::newton_krylov / chebyshev_inexact  krylov_newton.hpp:69-189
//   numa::hip::tf_chebyshev_2018 / tensor_free_curvature_2018
//                               methods/tf_chebyshev_2018.hpp:23-132
//   numa::hip::EigenEstimate / inverse_smallest_eigenpair
//                               analysis/spectral.hpp:22-115 (apply = J(x) v,
//                               precond = the provider's preconditioner)
//   numa::hip::gmres            linalg/gmres.hpp:61-65 (apply = J(x0) v)
// The solves run entirely in HBM; host vectors are staged once per call.
// Misuse throws std::invalid_argument / std::overflow_error as the reference
// does; numerical failure is reported in the result structs.

#include <cstddef>
#include <memory>
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

// problems/bratu3d.hpp:12-100 on the device.  eval / jvp / d2 stage host
// vectors; the solvers below never do.
class Bratu3D {
public:
    Bratu3D(const Context& c, std::size_t n, double lambda = 1.0) : c_(&c), n_(n), lambda_(lambda) {
        if (n == 0) throw std::invalid_argument("Bratu3D needs n > 0");
    }
    std::size_t size() const { return n_ * n_ * n_; }
    std::size_t grid() const { return n_; }
    double lambda() const { return lambda_; }
    const Context& context() const { return *c_; }

    void eval(const DVec& u, DVec& out) const {
        DeviceVec du(*c_, size()), dF(*c_, size());
        du.upload(u);
        check(numa_hip_bratu_eval(c_->get(), (int64_t)n_, lambda_, du.data(), dF.data(), nullptr));
        dF.download(out);
    }
    void jvp(const DVec& u, const DVec& v, DVec& out) const {
        binary(u, v, out, numa_hip_bratu_jvp);
    }
    void d2(const DVec& u, const DVec& v, DVec& out) const { binary(u, v, out, numa_hip_bratu_d2); }
    void jacobi_diag(const DVec& u, DVec& d) const {
        DeviceVec du(*c_, size()), dd(*c_, size());
        du.upload(u);
        check(numa_hip_bratu_jacobi_diag(c_->get(), (int64_t)n_, lambda_, du.data(), dd.data()));
        dd.download(d);
    }

private:
    template <class Fn>
    void binary(const DVec& u, const DVec& v, DVec& out, Fn fn) const {
        DeviceVec du(*c_, size()), dv(*c_, size()), dout(*c_, size());
        du.upload(u);
        dv.upload(v);
        check(fn(c_->get(), (int64_t)n_, lambda_, du.data(), dv.data(), dout.data()));
        dout.download(out);
    }
    const Context* c_;
    std::size_t n_;
    double lambda_;
};

static_assert(numa::op::SecondOrderOperator<Bratu3D>);

// Preconditioner providers: configuration only; the device solver binds them
// to the Newton point once per outer iteration as Provider::make does.
struct NoPrecond {
    numa_solver_config apply(numa_solver_config c) const {
        c.precond = NUMA_PRECOND_NONE;
        return c;
    }
};

struct BratuJacobi {
    const Bratu3D& f;
    numa_solver_config apply(numa_solver_config c) const {
        c.precond = NUMA_PRECOND_JACOBI;
        return c;
    }
};

struct PoissonMGProvider {
    numa_mg_config mg;
    // PoissonMG(n): RB-GS V(2,2), coarsening while n is odd (poisson_mg.hpp:26-64)
    explicit PoissonMGProvider(std::size_t = 0)
        : mg{NUMA_SMOOTH_RBGS, 2, 2, 3, 0, 1, 30, 0, 6.0 / 7.0, 4.0} {}
    explicit PoissonMGProvider(const numa_mg_config& cfg) : mg(cfg) {}
    numa_solver_config apply(numa_solver_config c) const {
        c.precond = NUMA_PRECOND_GMG;
        c.mg = mg;
        return c;
    }
};

namespace detail {

inline numa_solver_config base_config(int method, const SolverOptions& o) {
    numa_solver_config c{};
    c.method = method;
    c.precision = o.precision;
    c.flexible = o.flexible ? 1 : 0;
    c.jv_mode = o.jv_mode;
    c.d2_mode = o.d2_mode;
    c.mg = PoissonMGProvider().mg;
    return c;
}

inline KrylovResult solve(const Bratu3D& f, DVec x, const KrylovParams& p,
                          const numa_solver_config& cfg, Timers* timers = nullptr) {
    if (x.size() != f.size()) throw std::invalid_argument("x has the wrong size");
    numa_hip_solver* s = nullptr;
    check(numa_hip_bratu_solver_create(f.context().get(), (int64_t)f.grid(), f.lambda(), &cfg,
                                       (int64_t)std::max<std::size_t>(1, p.gmres_restart), &s));
    std::unique_ptr<numa_hip_solver, int (*)(numa_hip_solver*)> guard(s, numa_hip_solver_destroy);
    const numa_krylov_params cp = p.c();
    numa_krylov_result r{};
    if (timers) check(numa_hip_solver_set_timing(s, 1));
    check(numa_hip_solve_host(s, &cp, x.data(), &r));
    if (timers) {
        numa_timers t{};
        check(numa_hip_solver_timers(s, &t, 0));
        timers->t_F += t.t_F;
        timers->t_JVP += t.t_JVP;
        timers->t_D2 += t.t_D2;
        timers->t_prec += t.t_prec;
        timers->t_setup_dynamic += t.t_setup_dynamic;
    }
    KrylovResult out;
    out.x = std::move(x);
    out.converged = r.converged != 0;
    out.iters = (std::size_t)r.iters;
    out.n_F = (std::size_t)r.n_F;
    out.n_JVP = (std::size_t)r.n_JVP;
    out.n_D2 = (std::size_t)r.n_D2;
    out.n_lin_iters = (std::size_t)r.n_lin_iters;
    out.n_lin_failures = (std::size_t)r.n_lin_failures;
    out.n_prec = (std::size_t)r.n_prec;
    out.res0 = r.res0;
    out.res = r.res;
    out.t_total = r.t_total;
    return out;
}

}  // namespace detail

// methods/krylov_newton.hpp:69-71
template <class PrecondProvider = NoPrecond>
KrylovResult newton_krylov(const Bratu3D& f, DVec x, const KrylovParams& p = {},
                           const PrecondProvider& pc = {}, const SolverOptions& o = {}) {
    return detail::solve(f, std::move(x), p, pc.apply(detail::base_config(NUMA_METHOD_NEWTON, o)),
                         o.timers);
}

// methods/krylov_newton.hpp:116-118
template <class PrecondProvider = NoPrecond>
KrylovResult chebyshev_inexact(const Bratu3D& f, DVec x, const KrylovParams& p = {},
                               const PrecondProvider& pc = {}, const SolverOptions& o = {}) {
    return detail::solve(f, std::move(x), p,
                         pc.apply(detail::base_config(NUMA_METHOD_CHEBYSHEV, o)), o.timers);
}

// methods/tf_chebyshev_2018.hpp:53-57
template <class PrecondProvider = NoPrecond>
KrylovResult tf_chebyshev_2018(const Bratu3D& f, DVec x, const KrylovParams& p = {},
                               const PrecondProvider& pc = {}, const SolverOptions& o = {}) {
    return detail::solve(f, std::move(x), p,
                         pc.apply(detail::base_config(NUMA_METHOD_TF2018, o)), o.timers);
}

// methods/tf_chebyshev_2018.hpp:31-51: out = [J(x + h d) d - J(x) d] / h with
// the operator's analytic J v (o.jv_mode selects the finite-difference one).
inline void tensor_free_curvature_2018(const Bratu3D& f, const DVec& x, const DVec& d1, DVec& out,
                                       double* h_used = nullptr, const SolverOptions& o = {}) {
    const std::size_t N = f.size();
    if (x.size() != N || d1.size() != N) throw std::invalid_argument("vector size mismatch");
    const numa_solver_config cfg = NoPrecond{}.apply(detail::base_config(NUMA_METHOD_TF2018, o));
    numa_hip_solver* s = nullptr;
    check(numa_hip_bratu_solver_create(f.context().get(), (int64_t)f.grid(), f.lambda(), &cfg, 1,
                                       &s));
    std::unique_ptr<numa_hip_solver, int (*)(numa_hip_solver*)> guard(s, numa_hip_solver_destroy);
    DeviceVec dx(f.context(), N), dd(f.context(), N), dout(f.context(), N);
    dx.upload(x);
    dd.upload(d1);
    check(numa_hip_tf2018_curvature(s, dx.data(), dd.data(), dout.data(), h_used));
    out.resize(N);
    dout.download(out);
}

// analysis/spectral.hpp:22-32
struct EigenEstimate {
    bool converged = false;
    double eigenvalue = 0.0;
    double residual_rel = 0.0;
    std::size_t outer_iterations = 0;
    std::size_t linear_iterations = 0;
    std::size_t linear_matvecs = 0;
    std::size_t preconditioner_applications = 0;
    std::size_t linear_failures = 0;
    DVec vector;
};

// analysis/spectral.hpp:34-42 with apply(v, out) = J(x) v (the operator's J v,
// or o.jv_mode's) and the provider's preconditioner; FP64.
template <class PrecondProvider = NoPrecond>
EigenEstimate inverse_smallest_eigenpair(const Bratu3D& f, const DVec& x,
                                         const PrecondProvider& pc = {},
                                         double eigen_rtol = 1e-7, double linear_rtol = 1e-10,
                                         std::size_t max_outer = 30, std::size_t restart = 100,
                                         std::size_t max_linear = 2000,
                                         const SolverOptions& o = {}) {
    const std::size_t N = f.size();
    if (x.size() != N) throw std::invalid_argument("x has the wrong size");
    numa_solver_config cfg = pc.apply(detail::base_config(NUMA_METHOD_NEWTON, o));
    cfg.precision = NUMA_PRECISION_FP64;
    cfg.flexible = 0;
    numa_hip_solver* s = nullptr;
    check(numa_hip_bratu_solver_create(f.context().get(), (int64_t)f.grid(), f.lambda(), &cfg,
                                       (int64_t)std::max<std::size_t>(1, restart), &s));
    std::unique_ptr<numa_hip_solver, int (*)(numa_hip_solver*)> guard(s, numa_hip_solver_destroy);
    DeviceVec dx(f.context(), N), dv(f.context(), N);
    dx.upload(x);
    const numa_eigen_params p{eigen_rtol, linear_rtol, (int64_t)max_outer, (int64_t)restart,
                              (int64_t)max_linear};
    numa_eigen_result r{};
    check(numa_hip_eigen_smallest(s, dx.data(), &p, &r, dv.data()));
    EigenEstimate e;
    e.converged = r.converged != 0;
    e.eigenvalue = r.eigenvalue;
    e.residual_rel = r.residual_rel;
    e.outer_iterations = (std::size_t)r.outer_iterations;
    e.linear_iterations = (std::size_t)r.linear_iterations;
    e.linear_matvecs = (std::size_t)r.linear_matvecs;
    e.preconditioner_applications = (std::size_t)r.preconditioner_applications;
    e.linear_failures = (std::size_t)r.linear_failures;
    if (r.outer_iterations > 0) dv.download(e.vector);
    return e;
}

// linalg/gmres.hpp:61-65 with apply(v, out) = J(x0) v and the provider's
// preconditioner bound at x0.
template <class PrecondProvider = NoPrecond>
GmresResult gmres(const Bratu3D& f, const DVec& x0, const PrecondProvider& pc, const DVec& b,
                  DVec& x, double rtol, std::size_t restart, std::size_t max_iters,
                  bool zero_guess = true, const SolverOptions& o = {}) {
    const numa_solver_config cfg = pc.apply(detail::base_config(NUMA_METHOD_NEWTON, o));
    numa_hip_solver* s = nullptr;
    check(numa_hip_bratu_solver_create(f.context().get(), (int64_t)f.grid(), f.lambda(), &cfg,
                                       (int64_t)std::max<std::size_t>(1, restart), &s));
    std::unique_ptr<numa_hip_solver, int (*)(numa_hip_solver*)> guard(s, numa_hip_solver_destroy);
    const std::size_t N = f.size();
    if (x0.size() != N || b.size() != N) throw std::invalid_argument("vector size mismatch");
    if (x.size() != N) x.assign(N, 0.0);
    DeviceVec dx0(f.context(), N), db(f.context(), N), dz(f.context(), N);
    dx0.upload(x0);
    db.upload(b);
    dz.upload(x);
    numa_gmres_result r{};
    check(numa_hip_gmres(s, dx0.data(), db.data(), dz.data(), rtol, (int64_t)restart,
                         (int64_t)max_iters, zero_guess ? 1 : 0, &r));
    dz.download(x);
    return GmresResult{(std::size_t)r.iters, (std::size_t)r.matvecs, (std::size_t)r.precs,
                       r.rel_res, r.converged != 0, r.breakdown != 0};
}

}  // namespace numa::hip
