// ref_driver.cpp - C ABI glue around the reference's own header-only C++20
// library (kapumota/chebyshev-nonlinear-lab, include/numa), compiled from the
// headers where they lie by oracle/Makefile into oracle/_ref/libnuma_ref.so.
//
// TEST INFRASTRUCTURE ONLY.  Used in this container to pin the CPU oracle
// (oracle/numa_oracle.cpp) against the reference's own arithmetic and to write
// the fixtures under tests/golden/.  Nothing here is a product path.

#include <numa/analysis/spectral.hpp>
#include <numa/functions/gauss_legendre_dynamic.hpp>
#include <numa/linalg/gmres.hpp>
#include <numa/linalg/poisson_mg.hpp>
#include <numa/methods/krylov_newton.hpp>
#include <numa/methods/tf_chebyshev_2018.hpp>
#include <numa/operators/operator.hpp>
#include <numa/problems/bratu3d.hpp>
#include <numa/problems/chandrasekhar_op.hpp>

#include <cstdint>
#include <cstring>
#include <limits>

using namespace numa::op;

namespace {

DVec vec(const double* p, int64_t n) { return DVec(p, p + n); }
void put(const DVec& v, double* p) { std::memcpy(p, v.data(), v.size() * sizeof(double)); }

// Bratu3D with the generic finite-difference derivatives of operator.hpp
// (jvp_fd :101-114, d2_fd :118-134) in place of the analytic ones.
struct FDBratu {
    const Bratu3D& f;
    std::size_t size() const { return f.size(); }
    void eval(const DVec& x, DVec& out) const { f.eval(x, out); }
    void jvp(const DVec& x, const DVec& v, DVec& out) const { jvp_fd(f, x, v, out); }
    void d2(const DVec& x, const DVec& v, DVec& out) const { d2_fd(f, x, v, out); }
};

// Bratu3D (analytic derivatives) whose `at`-th eval (what 0) or jvp (what 1)
// returns NaN in element 0, once: the failure paths of the reference's own
// loops under a non-finite F or J v.
struct NanBratu {
    const Bratu3D& f;
    int what;
    std::size_t at;
    mutable std::size_t calls = 0;
    std::size_t size() const { return f.size(); }
    void poison(int w, DVec& out) const {
        if (w == what && ++calls == at) out[0] = std::numeric_limits<double>::quiet_NaN();
    }
    void eval(const DVec& x, DVec& out) const {
        f.eval(x, out);
        poison(0, out);
    }
    void jvp(const DVec& x, const DVec& v, DVec& out) const {
        f.jvp(x, v, out);
        poison(1, out);
    }
    void d2(const DVec& x, const DVec& v, DVec& out) const { f.d2(x, v, out); }
};

KrylovParams to_params(const double* p) {
    // p: rtol_nl, max_nl, eta_init, eta_max, gamma, alpha, eta2_factor,
    //    eta2_max, restart, max_iters, armijo, min_lambda, recycle
    KrylovParams k;
    k.rtol_nl = p[0];
    k.max_nl = (std::size_t)p[1];
    k.eta_init = p[2];
    k.eta_max = p[3];
    k.gamma = p[4];
    k.alpha = p[5];
    k.eta2_factor = p[6];
    k.eta2_max = p[7];
    k.gmres_restart = (std::size_t)p[8];
    k.gmres_max_iters = (std::size_t)p[9];
    k.armijo = p[10];
    k.min_lambda = p[11];
    k.recycle = p[12] != 0.0;
    return k;
}

void from_result(const KrylovResult& r, double* o) {
    // o: converged, iters, n_F, n_JVP, n_D2, n_lin_iters, n_lin_failures, n_prec, res0, res
    o[0] = r.converged;
    o[1] = (double)r.iters;
    o[2] = (double)r.n_F;
    o[3] = (double)r.n_JVP;
    o[4] = (double)r.n_D2;
    o[5] = (double)r.n_lin_iters;
    o[6] = (double)r.n_lin_failures;
    o[7] = (double)r.n_prec;
    o[8] = r.res0;
    o[9] = r.res;
}

template <class Op, class Prov>
KrylovResult run(const Op& f, const DVec& x0, const KrylovParams& p, const Prov& pc, int method) {
    if (method == 2) return tf_chebyshev_2018(f, x0, p, pc);
    return method == 1 ? chebyshev_inexact(f, x0, p, pc) : newton_krylov(f, x0, p, pc);
}

template <class Op>
KrylovResult run_prec(const Op& f, const Bratu3D& base, const DVec& x0, const KrylovParams& p,
                      int method, int precond) {
    if (precond == 1) return run(f, x0, p, BratuJacobi{base}, method);
    if (precond == 2) return run(f, x0, p, PoissonMGProvider(base.grid()), method);
    return run(f, x0, p, NoPrecond{}, method);
}

}  // namespace

extern "C" {

void ref_bratu_eval(int64_t n, double lambda, const double* u, double* out) {
    Bratu3D f((std::size_t)n, lambda);
    DVec o;
    f.eval(vec(u, n * n * n), o);
    put(o, out);
}
void ref_bratu_jvp(int64_t n, double lambda, const double* u, const double* v, double* out) {
    Bratu3D f((std::size_t)n, lambda);
    DVec o;
    f.jvp(vec(u, n * n * n), vec(v, n * n * n), o);
    put(o, out);
}
void ref_bratu_d2(int64_t n, double lambda, const double* u, const double* v, double* out) {
    Bratu3D f((std::size_t)n, lambda);
    DVec o;
    f.d2(vec(u, n * n * n), vec(v, n * n * n), o);
    put(o, out);
}
void ref_bratu_jacobi_diag(int64_t n, double lambda, const double* u, double* out) {
    Bratu3D f((std::size_t)n, lambda);
    DVec o;
    f.jacobi_diag(vec(u, n * n * n), o);
    put(o, out);
}
void ref_bratu_jvp_fd(int64_t n, double lambda, const double* x, const double* v, double* out) {
    Bratu3D f((std::size_t)n, lambda);
    DVec o;
    jvp_fd(f, vec(x, n * n * n), vec(v, n * n * n), o);
    put(o, out);
}
void ref_bratu_d2_fd(int64_t n, double lambda, const double* x, const double* v, double* out) {
    Bratu3D f((std::size_t)n, lambda);
    DVec o;
    d2_fd(f, vec(x, n * n * n), vec(v, n * n * n), o);
    put(o, out);
}

int64_t ref_mg_levels(int64_t n) { return (int64_t)PoissonMG((std::size_t)n).levels(); }

void ref_mg_apply(int64_t n, const double* r, double* z) {
    PoissonMG mg((std::size_t)n);
    DVec o;
    mg(vec(r, n * n * n), o);
    put(o, z);
}

// GMRES on J(x0) (analytic jvp if fd == 0, jvp_fd otherwise), preconditioner
// 0 none / 1 BratuJacobi / 2 PoissonMG.  res: iters, matvecs, precs, rel_res,
// converged, breakdown.
void ref_bratu_gmres(int64_t n, double lambda, int fd, int precond, const double* x0,
                     const double* b, double* x, double rtol, int64_t restart, int64_t max_iters,
                     double* res) {
    Bratu3D f((std::size_t)n, lambda);
    const int64_t N = n * n * n;
    const DVec X0 = vec(x0, N), B = vec(b, N);
    DVec X = vec(x, N);
    auto apply = [&](const DVec& v, DVec& o) {
        if (fd)
            jvp_fd(f, X0, v, o);
        else
            f.jvp(X0, v, o);
    };
    GmresResult g;
    if (precond == 1) {
        g = gmres(apply, BratuJacobi{f}.make(X0), B, X, rtol, (std::size_t)restart,
                  (std::size_t)max_iters);
    } else if (precond == 2) {
        PoissonMG mg((std::size_t)n);
        g = gmres(apply, mg, B, X, rtol, (std::size_t)restart, (std::size_t)max_iters);
    } else {
        g = gmres(apply, IdentityPrecond{}, B, X, rtol, (std::size_t)restart,
                  (std::size_t)max_iters);
    }
    put(X, x);
    res[0] = (double)g.iters;
    res[1] = (double)g.matvecs;
    res[2] = (double)g.precs;
    res[3] = g.rel_res;
    res[4] = g.converged;
    res[5] = g.breakdown;
}

// inverse_smallest_eigenpair of the analytic J(x) preconditioned by
// PoissonMG(n), as calibrate_b6.cpp:240-248 calls it.  res: converged,
// eigenvalue, residual_rel, outer, linear iterations, matvecs, precs, failures.
void ref_bratu_eigen_smallest(int64_t n, double lambda, const double* x, double eigen_rtol,
                              double linear_rtol, int64_t max_outer, int64_t restart,
                              int64_t max_linear, double* res, double* vec_out) {
    Bratu3D f((std::size_t)n, lambda);
    const DVec X = vec(x, n * n * n);
    PoissonMG mg((std::size_t)n);
    auto apply = [&](const DVec& v, DVec& o) { f.jvp(X, v, o); };
    const auto e = numa::analysis::inverse_smallest_eigenpair(
        apply, mg, f.size(), eigen_rtol, linear_rtol, (std::size_t)max_outer,
        (std::size_t)restart, (std::size_t)max_linear);
    res[0] = e.converged;
    res[1] = e.eigenvalue;
    res[2] = e.residual_rel;
    res[3] = (double)e.outer_iterations;
    res[4] = (double)e.linear_iterations;
    res[5] = (double)e.linear_matvecs;
    res[6] = (double)e.preconditioner_applications;
    res[7] = (double)e.linear_failures;
    if (e.vector.size() == f.size()) put(e.vector, vec_out);
}

// tensor_free_curvature_2018 with the analytic (fd = 0) or FD (fd = 1) J v.
void ref_tf2018_curvature(int64_t n, double lambda, int fd, const double* x, const double* d,
                          double* out, double* h_used) {
    Bratu3D f((std::size_t)n, lambda);
    const int64_t N = n * n * n;
    DVec o;
    if (fd)
        tensor_free_curvature_2018(FDBratu{f}, vec(x, N), vec(d, N), o, h_used);
    else
        tensor_free_curvature_2018(f, vec(x, N), vec(d, N), o, h_used);
    put(o, out);
}

// method 0 newton_krylov / 1 chebyshev_inexact / 2 tf_chebyshev_2018;
// fd: 0 analytic, 1 FD Jv and F''.
void ref_bratu_solve(int64_t n, double lambda, int method, int precond, int fd,
                     const double* params, double* x, double* result) {
    Bratu3D f((std::size_t)n, lambda);
    const KrylovParams p = to_params(params);
    const DVec x0 = vec(x, n * n * n);
    KrylovResult r = fd ? run_prec(FDBratu{f}, f, x0, p, method, precond)
                        : run_prec(f, f, x0, p, method, precond);
    put(r.x, x);
    from_result(r, result);
}

// ref_bratu_solve with the analytic derivatives and a NaN injected into the
// `at`-th F evaluation (what 0) or J v product (what 1).
void ref_bratu_solve_inject(int64_t n, double lambda, int method, int precond,
                            const double* params, double* x, double* result, int what,
                            int64_t at) {
    Bratu3D f((std::size_t)n, lambda);
    const KrylovParams p = to_params(params);
    const DVec x0 = vec(x, n * n * n);
    const NanBratu g{f, what, (std::size_t)at};
    KrylovResult r = run_prec(g, f, x0, p, method, precond);
    put(r.x, x);
    from_result(r, result);
}

void ref_gauss_legendre_unit(int64_t n, double* nodes, double* weights) {
    const auto rule = numa::quadrature::gauss_legendre_unit((std::size_t)n);
    put(rule.nodes, nodes);
    put(rule.weights, weights);
}

void ref_chandra_eval(int64_t n, double c, const double* x, double* out) {
    ChandrasekharOp op((std::size_t)n, c);
    DVec o;
    op.eval(vec(x, n), o);
    put(o, out);
}
void ref_chandra_jvp(int64_t n, double c, const double* x, const double* v, double* out) {
    ChandrasekharOp op((std::size_t)n, c);
    DVec o;
    op.jvp(vec(x, n), vec(v, n), o);
    put(o, out);
}
void ref_chandra_d2(int64_t n, double c, const double* x, const double* v, double* out) {
    ChandrasekharOp op((std::size_t)n, c);
    DVec o;
    op.d2(vec(x, n), vec(v, n), o);
    put(o, out);
}
void ref_chandra_solve(int64_t n, double c, int method, const double* params, double* x,
                       double* result) {
    ChandrasekharOp op((std::size_t)n, c);
    const KrylovParams p = to_params(params);
    KrylovResult r = method == 1 ? chebyshev_inexact(op, vec(x, n), p)
                                 : newton_krylov(op, vec(x, n), p);
    put(r.x, x);
    from_result(r, result);
}

}  // extern "C"
