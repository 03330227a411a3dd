#pragma once
// problems/chandrasekhar_op.hpp:31-136 on the device:
//   F(x)_i = 1 + x_i (A x)_i - x_i,   (A v)_i = sum_j (c/2) w_j mu_i/(mu_i+mu_j) v_j
// with the Gauss-Legendre rule on [0, 1] (functions/gauss_legendre_dynamic.hpp),
// as a SecondOrderOperator (eval / jvp / d2 stage host vectors; the contraction
// runs on the FP64 matrix cores), and newton_krylov / chebyshev_inexact
// overloads that run the whole solve on the device (the batched Chandrasekhar
// solver with a batch of one).  The constructor rejects N = 0 and c outside
// [0, 1] with std::invalid_argument, and eval / jvp / d2 a vector of another
// length, as the reference does.

#include <cmath>

#include "core.hpp"

namespace numa::hip {

class ChandrasekharOp {
public:
    ChandrasekharOp(std::size_t n, double c) : ChandrasekharOp(default_context(), n, c) {}
    ChandrasekharOp(const Context& ctx, std::size_t n, double c) : ctx_(&ctx), n_(n), c_(c) {
        if (n_ == 0) throw std::invalid_argument("ChandrasekharOp needs N > 0");
        if (!std::isfinite(c_) || c_ < 0.0 || c_ > 1.0)
            throw std::invalid_argument("ChandrasekharOp needs a finite c in [0, 1]");
        mu_.resize(n_);
        w_.resize(n_);
        check(numa_hip_chandra_quadrature((int64_t)n_, 0, mu_.data(), w_.data()));
        h_ = make_handle(1);
    }
    ~ChandrasekharOp() {
        if (h_) numa_hip_chandra_destroy(h_);
    }
    ChandrasekharOp(const ChandrasekharOp&) = delete;
    ChandrasekharOp& operator=(const ChandrasekharOp&) = delete;

    std::size_t size() const { return n_; }
    double c() const { return c_; }
    const std::vector<double>& nodes() const { return mu_; }
    const std::vector<double>& weights() const { return w_; }
    const Context& context() const { return *ctx_; }

    void eval(const DVec& x, DVec& out) const {
        validate(x);
        DeviceVec dx(*ctx_, n_), dF(*ctx_, n_);
        dx.upload(x);
        check(numa_hip_chandra_eval(h_, dx.data(), dF.data(), nullptr));
        dF.download(out);
    }
    void jvp(const DVec& x, const DVec& v, DVec& out) const {
        binary(x, v, out, numa_hip_chandra_jvp);
    }
    void d2(const DVec& x, const DVec& v, DVec& out) const {
        binary(x, v, out, numa_hip_chandra_d2);
    }

    // A device handle of this operator for a solve (method NUMA_METHOD_*),
    // with a Krylov basis of `restart` vectors.
    numa_hip_chandra* make_handle(std::size_t restart, int method = NUMA_METHOD_NEWTON) const {
        numa_hip_chandra* h = nullptr;
        check(numa_hip_chandra_create(ctx_->get(), (int64_t)n_, 1, &c_, 0, 0,
                                      NUMA_PRECISION_FP64,
                                      (int64_t)std::max<std::size_t>(1, restart), &h));
        if (method != NUMA_METHOD_NEWTON) {
            const int rc = numa_hip_chandra_set_method(h, method);
            if (rc != NUMA_OK) {
                numa_hip_chandra_destroy(h);
                check(rc);
            }
        }
        return h;
    }

private:
    void validate(const DVec& v) const {
        if (v.size() != n_) throw std::invalid_argument("ChandrasekharOp: dimension mismatch");
    }
    template <class Fn>
    void binary(const DVec& x, const DVec& v, DVec& out, Fn fn) const {
        validate(x);
        validate(v);
        DeviceVec dx(*ctx_, n_), dv(*ctx_, n_), dout(*ctx_, n_);
        dx.upload(x);
        dv.upload(v);
        check(fn(h_, dx.data(), dv.data(), dout.data()));
        dout.download(out);
    }
    const Context* ctx_;
    std::size_t n_;
    double c_;
    std::vector<double> mu_, w_;
    numa_hip_chandra* h_ = nullptr;
};

static_assert(numa::op::SecondOrderOperator<ChandrasekharOp>);

namespace detail {

inline KrylovResult chandra_solve(const ChandrasekharOp& f, DVec x, const KrylovParams& p,
                                  int method) {
    if (x.size() != f.size()) throw std::invalid_argument("x has the wrong size");
    std::unique_ptr<numa_hip_chandra, int (*)(numa_hip_chandra*)> h(
        f.make_handle(p.gmres_restart, method), numa_hip_chandra_destroy);
    DeviceVec dx(f.context(), f.size());
    dx.upload(x);
    const numa_krylov_params cp = p.c();
    numa_krylov_result r{};
    check(numa_hip_chandra_solve(h.get(), &cp, dx.data(), &r));
    KrylovResult out;
    dx.download(out.x);
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

// methods/krylov_newton.hpp:69-71 (no preconditioner: the reference's
// Chandrasekhar runs use NoPrecond)
inline KrylovResult newton_krylov(const ChandrasekharOp& f, DVec x, const KrylovParams& p = {}) {
    return detail::chandra_solve(f, std::move(x), p, NUMA_METHOD_NEWTON);
}

// methods/krylov_newton.hpp:116-118
inline KrylovResult chebyshev_inexact(const ChandrasekharOp& f, DVec x,
                                      const KrylovParams& p = {}) {
    return detail::chandra_solve(f, std::move(x), p, NUMA_METHOD_CHEBYSHEV);
}

// Forward difference (F(x + h v) - F(x)) / h, h = sqrt(eps)(1 + ||x||)/||v||,
// and 0 for v = 0, for any operator (operators/operator.hpp:101-114).
template <numa::op::NonlinearOperator Op>
void jvp_fd(const Op& f, const DVec& x, const DVec& v, DVec& out) {
    const std::size_t n = x.size();
    auto nrm = [](const DVec& a) {
        double s = 0.0;
        for (double e : a) s += e * e;
        return std::sqrt(s);
    };
    out.assign(n, 0.0);
    const double nv = nrm(v);
    if (nv == 0.0) return;
    const double h = std::sqrt(std::numeric_limits<double>::epsilon()) * (1.0 + nrm(x)) / nv;
    DVec xh(x), f0, f1;
    for (std::size_t i = 0; i < n; ++i) xh[i] += h * v[i];
    f.eval(x, f0);
    f.eval(xh, f1);
    for (std::size_t i = 0; i < n; ++i) out[i] = (f1[i] - f0[i]) / h;
}

// Central second difference (F(x + h v) - 2 F(x) + F(x - h v)) / h^2,
// h = eps^(1/4)(1 + ||x||)/||v|| (operators/operator.hpp:118-134).
template <numa::op::NonlinearOperator Op>
void d2_fd(const Op& f, const DVec& x, const DVec& v, DVec& out) {
    const std::size_t n = x.size();
    auto nrm = [](const DVec& a) {
        double s = 0.0;
        for (double e : a) s += e * e;
        return std::sqrt(s);
    };
    out.assign(n, 0.0);
    const double nv = nrm(v);
    if (nv == 0.0) return;
    const double h = std::pow(std::numeric_limits<double>::epsilon(), 0.25) * (1.0 + nrm(x)) / nv;
    DVec xp(x), xm(x), f0, fp, fm;
    for (std::size_t i = 0; i < n; ++i) {
        xp[i] += h * v[i];
        xm[i] -= h * v[i];
    }
    f.eval(x, f0);
    f.eval(xp, fp);
    f.eval(xm, fm);
    for (std::size_t i = 0; i < n; ++i) out[i] = (fp[i] - 2.0 * f0[i] + fm[i]) / (h * h);
}

}  // namespace numa::hip
