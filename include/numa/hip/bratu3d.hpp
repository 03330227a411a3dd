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

#include "core.hpp"
#include "gmres.hpp"

namespace numa::hip {

// problems/bratu3d.hpp:12-100 on the device.  eval / jvp / d2 stage host
// vectors; the solvers below never do.
class Bratu3D {
public:
    Bratu3D(const Context& c, std::size_t n, double lambda = 1.0) : c_(&c), n_(n), lambda_(lambda) {
        if (n == 0) throw std::invalid_argument("Bratu3D needs n > 0");
    }
    // the reference's signature, on default_context()
    explicit Bratu3D(std::size_t n, double lambda = 1.0) : Bratu3D(default_context(), n, lambda) {}
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

// PoissonMG (linalg/poisson_mg.hpp:24-299): the V-cycle for -Lap_h on n^3
// nodes as a preconditioner object, operator()(r, z) (poisson_mg.hpp:69-80).
// The default configuration is the reference's (RB-GS V(2,2), coarsening
// while n is odd); a numa_mg_config selects the even-n hierarchy and the
// weighted-Jacobi / Chebyshev smoothers.  Not thread safe, as the reference.
class PoissonMG {
public:
    static numa_mg_config reference_config() {
        return numa_mg_config{NUMA_SMOOTH_RBGS, 2, 2, 3, 0, 1, 30, 0, 6.0 / 7.0, 4.0, 0, 0, 0};
    }
    explicit PoissonMG(std::size_t n) : PoissonMG(default_context(), n, reference_config()) {}
    PoissonMG(const Context& c, std::size_t n, const numa_mg_config& cfg = reference_config())
        : c_(&c), n_(n) {
        if (n == 0) throw std::invalid_argument("PoissonMG needs n > 0");
        check(numa_hip_mg_create(c.get(), (int64_t)n, &cfg, NUMA_PRECISION_FP64, &h_));
    }
    ~PoissonMG() {
        if (h_) numa_hip_mg_destroy(h_);
    }
    PoissonMG(const PoissonMG&) = delete;
    PoissonMG& operator=(const PoissonMG&) = delete;
    std::size_t levels() const {
        int64_t l = 0;
        check(numa_hip_mg_levels(h_, &l, nullptr));
        return (std::size_t)l;
    }
    std::size_t size() const { return n_ * n_ * n_; }
    void operator()(const DVec& r, DVec& z) const {
        if (r.size() != size()) throw std::invalid_argument("PoissonMG: dimension mismatch");
        DeviceVec dr(*c_, size()), dz(*c_, size());
        dr.upload(r);
        check(numa_hip_mg_apply(h_, dr.data(), dz.data()));
        dz.download(z);
    }

private:
    const Context* c_;
    std::size_t n_;
    numa_hip_mg* h_ = nullptr;
};

// Preconditioner providers (krylov_newton.hpp:20-22).  make(x) binds the
// preconditioner at a point and returns it as a callable (r, z) -> z = M^-1 r
// for code that drives GMRES itself (bratu3d.hpp:104-120,
// poisson_mg.hpp:303-320); apply(config) is what the device solvers below
// use, which bind it on the device once per outer iteration instead.
struct NoPrecond {
    IdentityPrecond make(const DVec&) const { return {}; }
    numa_solver_config apply(numa_solver_config c) const {
        c.precond = NUMA_PRECOND_NONE;
        return c;
    }
};

struct BratuJacobi {
    const Bratu3D& f;
    // z = r / (6 (n+1)^2 - lambda e^x), the diagonal computed on the device
    struct Bound {
        DVec d;
        void operator()(const DVec& r, DVec& z) const {
            z.resize(r.size());
            for (std::size_t i = 0; i < r.size(); ++i) z[i] = r[i] / d[i];
        }
    };
    Bound make(const DVec& x) const {
        Bound b;
        f.jacobi_diag(x, b.d);
        return b;
    }
    numa_solver_config apply(numa_solver_config c) const {
        c.precond = NUMA_PRECOND_JACOBI;
        return c;
    }
};

struct PoissonMGProvider {
    numa_mg_config mg;
    std::size_t n = 0;
    const Context* ctx = nullptr;
    // PoissonMG(n): RB-GS V(2,2), coarsening while n is odd (poisson_mg.hpp:26-64)
    explicit PoissonMGProvider(std::size_t n_ = 0)
        : mg(PoissonMG::reference_config()), n(n_) {}
    explicit PoissonMGProvider(const numa_mg_config& cfg, std::size_t n_ = 0) : mg(cfg), n(n_) {}
    PoissonMGProvider(const Context& c, std::size_t n_, const numa_mg_config& cfg)
        : mg(cfg), n(n_), ctx(&c) {}
    // The hierarchy is built once per provider and shared by every bound
    // copy (poisson_mg.hpp:304-319: make binds without set-up).
    struct Bound {
        std::shared_ptr<PoissonMG> m;
        void operator()(const DVec& r, DVec& z) const { (*m)(r, z); }
    };
    Bound make(const DVec&) const {
        if (n == 0) throw std::invalid_argument("PoissonMGProvider: grid size not given");
        if (!shared_)
            shared_ = std::make_shared<PoissonMG>(ctx ? *ctx : default_context(), n, mg);
        return Bound{shared_};
    }
    numa_solver_config apply(numa_solver_config c) const {
        c.precond = NUMA_PRECOND_GMG;
        c.mg = mg;
        return c;
    }

private:
    mutable std::shared_ptr<PoissonMG> shared_;
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

// The outer loops on a Counted<Bratu3D> (operators/operator.hpp:68-98): the
// solve runs on the wrapped operator's device kernels and the wrapper counts
// the F, J v and F'' calls the solver reports, as the reference's wrapper
// counts the calls made through it.
template <class PrecondProvider = NoPrecond>
KrylovResult newton_krylov(const numa::op::Counted<Bratu3D>& f, DVec x,
                           const KrylovParams& p = {}, const PrecondProvider& pc = {},
                           const SolverOptions& o = {}) {
    auto r = newton_krylov(f.base(), std::move(x), p, pc, o);
    f.add(numa::op::OpCounters{r.n_F, r.n_JVP, r.n_D2});
    return r;
}
template <class PrecondProvider = NoPrecond>
KrylovResult chebyshev_inexact(const numa::op::Counted<Bratu3D>& f, DVec x,
                               const KrylovParams& p = {}, const PrecondProvider& pc = {},
                               const SolverOptions& o = {}) {
    auto r = chebyshev_inexact(f.base(), std::move(x), p, pc, o);
    f.add(numa::op::OpCounters{r.n_F, r.n_JVP, r.n_D2});
    return r;
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

namespace detail {

inline GmresResult from_c(const numa_gmres_result& r) {
    return GmresResult{(std::size_t)r.iters, (std::size_t)r.matvecs, (std::size_t)r.precs,
                       r.rel_res, r.converged != 0, r.breakdown != 0};
}

// A device solver of J(x0) for the GMRES overloads below.
template <class PrecondProvider>
std::unique_ptr<numa_hip_solver, int (*)(numa_hip_solver*)> jacobian_solver(
    const Bratu3D& f, const PrecondProvider& pc, std::size_t restart, const SolverOptions& o) {
    const numa_solver_config cfg = pc.apply(base_config(NUMA_METHOD_NEWTON, o));
    numa_hip_solver* s = nullptr;
    check(numa_hip_bratu_solver_create(f.context().get(), (int64_t)f.grid(), f.lambda(), &cfg,
                                       (int64_t)std::max<std::size_t>(1, restart), &s));
    return {s, numa_hip_solver_destroy};
}

// Downloads the relation a device GMRES saved (V_i = s_i Vraw_i).
inline void export_space(numa_hip_solver* s, const Context& c, std::size_t N, KrylovSpace& out) {
    out.clear();
    int64_t k = 0;
    check(numa_hip_gmres_space(s, &k, nullptr, nullptr));
    if (k <= 0) return;
    std::vector<double> hb((size_t)((k + 1) * k));
    std::vector<std::unique_ptr<DeviceVec>> dv;
    std::vector<double*> ptrs;
    for (int64_t i = 0; i <= k; ++i) {
        dv.push_back(std::make_unique<DeviceVec>(c, N));
        ptrs.push_back(dv.back()->data());
    }
    check(numa_hip_gmres_space(s, &k, hb.data(), ptrs.data()));
    out.k = (std::size_t)k;
    out.V.resize((size_t)k + 1);
    for (int64_t i = 0; i <= k; ++i) dv[(size_t)i]->download(out.V[(size_t)i]);
    out.Hbar.assign((size_t)k + 1, std::vector<double>((size_t)k, 0.0));
    for (int64_t i = 0; i <= k; ++i)
        for (int64_t j = 0; j < k; ++j) out.Hbar[(size_t)i][(size_t)j] = hb[(size_t)(i * k + j)];
}

}  // namespace detail

// linalg/gmres.hpp:61-65 with apply(v, out) = J(x0) v and the provider's
// preconditioner bound at x0, all on the device: save receives the last
// cycle's relation (KrylovSpace, gmres.hpp:25-35), r_init replaces the
// initial residual b - J x.
template <class PrecondProvider = NoPrecond>
GmresResult gmres(const Bratu3D& f, const DVec& x0, const PrecondProvider& pc, const DVec& b,
                  DVec& x, double rtol, std::size_t restart, std::size_t max_iters,
                  bool zero_guess = true, KrylovSpace* save = nullptr,
                  const DVec* r_init = nullptr, const SolverOptions& o = {}) {
    const std::size_t N = f.size();
    if (x0.size() != N || b.size() != N) throw std::invalid_argument("vector size mismatch");
    if (r_init && r_init->size() != N) {  // gmres.hpp:86-90
        GmresResult r;
        r.rel_res = 1.0;
        r.breakdown = true;
        return r;
    }
    auto s = detail::jacobian_solver(f, pc, restart, o);
    if (x.size() != N) x.assign(N, 0.0);
    DeviceVec dx0(f.context(), N), db(f.context(), N), dz(f.context(), N);
    dx0.upload(x0);
    db.upload(b);
    dz.upload(x);
    std::unique_ptr<DeviceVec> dr;
    if (r_init) {
        dr = std::make_unique<DeviceVec>(f.context(), N);
        dr->upload(*r_init);
    }
    numa_gmres_result r{};
    check(numa_hip_gmres_ex(s.get(), dx0.data(), db.data(), dz.data(), rtol, (int64_t)restart,
                            (int64_t)max_iters, zero_guess ? 1 : 0, dr ? dr->data() : nullptr,
                            save ? 1 : 0, &r));
    dz.download(x);
    if (save) detail::export_space(s.get(), f.context(), N, *save);
    return detail::from_c(r);
}

// linalg/gmres.hpp:229-323 on J(x0), on the device: b projected on `space`
// without J v products, then GMRES from the projected residual.
template <class PrecondProvider = NoPrecond>
GmresResult gmres_recycled(const Bratu3D& f, const DVec& x0, const PrecondProvider& pc,
                           const DVec& b, DVec& x, double rtol, std::size_t restart,
                           std::size_t max_iters, const KrylovSpace& space,
                           KrylovSpace* save = nullptr, const SolverOptions& o = {}) {
    const std::size_t N = f.size();
    if (x0.size() != N || b.size() != N) throw std::invalid_argument("vector size mismatch");
    const bool usable = detail::usable_space(space, N) && space.k <= restart;
    auto s = detail::jacobian_solver(f, pc, std::max(restart, usable ? space.k : 0), o);
    DeviceVec dx0(f.context(), N), db(f.context(), N), dz(f.context(), N);
    dx0.upload(x0);
    db.upload(b);
    std::vector<std::unique_ptr<DeviceVec>> dv;
    std::vector<const double*> ptrs;
    std::vector<double> hb;
    const int64_t k = usable ? (int64_t)space.k : 0;
    for (int64_t i = 0; i <= k && k > 0; ++i) {
        dv.push_back(std::make_unique<DeviceVec>(f.context(), N));
        dv.back()->upload(space.V[(size_t)i]);
        ptrs.push_back(dv.back()->data());
        for (int64_t j = 0; j < k; ++j) hb.push_back(space.Hbar[(size_t)i][(size_t)j]);
    }
    if (k == 0) {  // no usable relation: the loader sees an empty one
        dv.push_back(std::make_unique<DeviceVec>(f.context(), N));
        ptrs.push_back(dv.back()->data());
    }
    numa_gmres_result r{};
    check(numa_hip_gmres_recycled(s.get(), dx0.data(), db.data(), dz.data(), rtol,
                                  (int64_t)restart, (int64_t)max_iters, k,
                                  hb.empty() ? nullptr : hb.data(), ptrs.data(), save ? 1 : 0,
                                  &r));
    x.resize(N);
    dz.download(x);
    if (save) detail::export_space(s.get(), f.context(), N, *save);
    return detail::from_c(r);
}

}  // namespace numa::hip
