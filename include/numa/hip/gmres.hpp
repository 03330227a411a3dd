#pragma once
// linalg/gmres.hpp (KrylovSpace :25-35, IdentityPrecond :37-39, gmres :61-223,
// gmres_recycled :229-323) for callers that hand GMRES their own operator and
// preconditioner as host callables, apply(v, out) = A v and precond(r, z) =
// M^-1 r on std::vector<double>.  The Krylov basis lives in HBM: the
// Gram-Schmidt dots and updates, the norms and the basis combinations run on
// the device (numa_hip_dot / axpy / scal); only each callable's operand and
// result cross PCIe.  The Givens rotations and the triangular solve are the
// small host part, as on the device solver.  Counters (iters, matvecs, precs),
// flags and the saved relation follow the reference's definitions.
//
// The Bratu3D Jacobian has its own device-resident overloads (bratu3d.hpp),
// which never stage a vector.

#include <cmath>

#include "core.hpp"

namespace numa::hip {

// linalg/gmres.hpp:25-35: the last cycle's relation A M^-1 V_k = V_k+1 Hbar.
struct KrylovSpace {
    std::vector<DVec> V;
    std::vector<std::vector<double>> Hbar;
    std::size_t k = 0;
    void clear() {
        V.clear();
        Hbar.clear();
        k = 0;
    }
};

// linalg/gmres.hpp:37-39
struct IdentityPrecond {
    void operator()(const DVec& r, DVec& z) const { z = r; }
};

namespace detail {

// A basis of device vectors of one length with the BLAS-1 the Arnoldi process
// needs.
class DeviceBasis {
public:
    DeviceBasis(const Context& c, std::size_t n) : c_(&c), n_(n) {}
    DeviceVec& vec(std::size_t i) {
        while (v_.size() <= i) v_.push_back(std::make_unique<DeviceVec>(*c_, n_));
        return *v_[i];
    }
    double dot(const DeviceVec& a, const DeviceVec& b) const {
        double r = 0.0;
        check(numa_hip_dot(c_->get(), (int64_t)n_, a.data(), b.data(), &r));
        return r;
    }
    void axpy(double a, const DeviceVec& x, DeviceVec& y) const {
        check(numa_hip_axpy(c_->get(), (int64_t)n_, a, x.data(), y.data()));
    }
    void scal(double a, const DeviceVec& x, DeviceVec& out) const {
        check(numa_hip_scal(c_->get(), (int64_t)n_, a, x.data(), out.data()));
    }

private:
    const Context* c_;
    std::size_t n_;
    std::vector<std::unique_ptr<DeviceVec>> v_;
};

inline bool usable_space(const KrylovSpace& s, std::size_t n) {
    if (s.k == 0 || s.V.size() < s.k + 1 || s.Hbar.size() < s.k + 1) return false;
    for (std::size_t i = 0; i <= s.k; ++i)
        if (s.Hbar[i].size() < s.k || s.V[i].size() != n) return false;
    return true;
}

}  // namespace detail

// linalg/gmres.hpp:61-65 with a host operator and preconditioner; `ctx` is the
// device that holds the basis.
template <class Apply, class Precond>
GmresResult gmres(const Apply& apply, const Precond& precond, const DVec& b, DVec& x,
                  double rtol, std::size_t restart, std::size_t max_iters,
                  bool zero_guess = true, KrylovSpace* save = nullptr,
                  const DVec* r_init = nullptr, const Context& ctx = default_context()) {
    const std::size_t n = b.size();
    GmresResult res;
    if (save) save->clear();
    detail::DeviceBasis B(ctx, n);
    DeviceVec& db = B.vec(0);  // slot 0: b; the basis starts at slot 1
    db.upload(b);
    const double bnorm = std::sqrt(B.dot(db, db));
    if (bnorm == 0.0) {
        x.assign(n, 0.0);
        res.converged = true;
        return res;
    }
    if (restart == 0 || (r_init && r_init->size() != n)) {
        res.rel_res = 1.0;
        res.breakdown = true;
        return res;
    }
    if (max_iters == 0) {
        res.rel_res = 1.0;
        return res;
    }
    if (zero_guess || x.size() != n) x.assign(n, 0.0);
    auto V = [&](std::size_t i) -> DeviceVec& { return B.vec(i + 1); };
    DeviceVec dw(ctx, n);
    std::vector<std::vector<double>> H(restart + 1, std::vector<double>(restart, 0.0)),
        Hsave(restart + 1, std::vector<double>(restart, 0.0));
    std::vector<double> cs(restart), sn(restart), g(restart + 1), y(restart);
    DVec hv(n), hz, hw;
    bool given = r_init != nullptr, from_b = zero_guess;
    std::size_t k_last = 0;
    while (res.iters < max_iters) {
        // cycle start: r = r_init, b, or b - A x
        if (given) {
            V(0).upload(*r_init);
            given = from_b = false;
        } else if (from_b) {
            B.scal(1.0, db, V(0));
            from_b = false;
        } else {
            apply(x, hw);
            ++res.matvecs;
            V(0).upload(hw);
            B.scal(-1.0, V(0), V(0));
            B.axpy(1.0, db, V(0));
        }
        const double beta = std::sqrt(B.dot(V(0), V(0)));
        res.rel_res = beta / bnorm;
        if (res.rel_res <= rtol) {
            res.converged = true;
            k_last = 0;
            break;
        }
        B.scal(1.0 / beta, V(0), V(0));
        for (auto& row : Hsave) std::fill(row.begin(), row.end(), 0.0);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        std::size_t k = 0;
        k_last = 0;
        while (k < restart && res.iters < max_iters) {
            const std::size_t j = k;
            V(j).download(hv);
            precond(hv, hz);
            ++res.precs;
            apply(hz, hw);
            ++res.matvecs;
            dw.upload(hw);
            for (std::size_t i = 0; i <= j; ++i) {  // modified Gram-Schmidt on the device
                H[i][j] = B.dot(dw, V(i));
                B.axpy(-H[i][j], V(i), dw);
            }
            H[j + 1][j] = std::sqrt(B.dot(dw, dw));
            B.scal(H[j + 1][j] > 0.0 ? 1.0 / H[j + 1][j] : 0.0, dw, V(j + 1));
            for (std::size_t i = 0; i <= j + 1; ++i) Hsave[i][j] = H[i][j];
            for (std::size_t i = 0; i < j; ++i) {  // earlier rotations
                const double t = cs[i] * H[i][j] + sn[i] * H[i + 1][j];
                H[i + 1][j] = -sn[i] * H[i][j] + cs[i] * H[i + 1][j];
                H[i][j] = t;
            }
            const double d = std::hypot(H[j][j], H[j + 1][j]);
            if (!(d > 0.0) || !std::isfinite(d)) {
                res.breakdown = true;
                return res;
            }
            cs[j] = H[j][j] / d;
            sn[j] = H[j + 1][j] / d;
            H[j][j] = d;
            H[j + 1][j] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            ++k;
            ++res.iters;
            k_last = k;
            res.rel_res = std::abs(g[j + 1]) / bnorm;
            if (res.rel_res <= rtol) break;
        }
        for (std::size_t ii = k; ii-- > 0;) {  // back substitution
            const double pivot = H[ii][ii];
            if (!(std::abs(pivot) > 0.0) || !std::isfinite(pivot)) {
                res.breakdown = true;
                return res;
            }
            double s = g[ii];
            for (std::size_t jj = ii + 1; jj < k; ++jj) s -= H[ii][jj] * y[jj];
            y[ii] = s / pivot;
        }
        B.scal(y[0], V(0), dw);  // dw = V y (k >= 1 here), then x += M^-1 (V y)
        for (std::size_t i = 1; i < k; ++i) B.axpy(y[i], V(i), dw);
        dw.download(hv);
        precond(hv, hz);
        ++res.precs;
        for (std::size_t i = 0; i < n; ++i) x[i] += hz[i];
        if (res.rel_res <= rtol) {
            res.converged = true;
            break;
        }
    }
    if (save && k_last > 0) {
        save->k = k_last;
        save->V.assign(restart + 1, DVec(n, 0.0));
        for (std::size_t i = 0; i <= restart; ++i) V(i).download(save->V[i]);
        save->Hbar.assign(k_last + 1, std::vector<double>(k_last, 0.0));
        for (std::size_t i = 0; i <= k_last; ++i)
            for (std::size_t j = 0; j < k_last; ++j) save->Hbar[i][j] = Hsave[i][j];
    }
    return res;
}

// linalg/gmres.hpp:229-323: b projected on the saved relation without A v
// products, x0 = M^-1 V_k y, then GMRES from the projected residual (or from
// zero when the space is unusable or the projection singular).
template <class Apply, class Precond>
GmresResult gmres_recycled(const Apply& apply, const Precond& precond, const DVec& b, DVec& x,
                           double rtol, std::size_t restart, std::size_t max_iters,
                           const KrylovSpace& space, KrylovSpace* save = nullptr,
                           const Context& ctx = default_context()) {
    const std::size_t n = b.size(), k = space.k;
    x.assign(n, 0.0);
    auto fresh = [&]() {
        x.assign(n, 0.0);
        return gmres(apply, precond, b, x, rtol, restart, max_iters, true, save, nullptr, ctx);
    };
    if (!detail::usable_space(space, n)) return fresh();
    detail::DeviceBasis B(ctx, n);
    B.vec(0).upload(b);
    for (std::size_t i = 0; i <= k; ++i) B.vec(i + 1).upload(space.V[i]);
    std::vector<double> rhs(k + 1);
    for (std::size_t i = 0; i <= k; ++i) rhs[i] = B.dot(B.vec(i + 1), B.vec(0));
    // least squares min || rhs - Hbar y || by Givens rotations
    auto R = space.Hbar;
    std::vector<double> cs(k), sn(k), y(k);
    for (std::size_t j = 0; j < k; ++j) {
        for (std::size_t i = 0; i < j; ++i) {
            const double t = cs[i] * R[i][j] + sn[i] * R[i + 1][j];
            R[i + 1][j] = -sn[i] * R[i][j] + cs[i] * R[i + 1][j];
            R[i][j] = t;
        }
        const double d = std::hypot(R[j][j], R[j + 1][j]);
        if (!(d > 0.0) || !std::isfinite(d)) return fresh();
        cs[j] = R[j][j] / d;
        sn[j] = R[j + 1][j] / d;
        R[j][j] = d;
        R[j + 1][j] = 0.0;
        const double t = cs[j] * rhs[j] + sn[j] * rhs[j + 1];
        rhs[j + 1] = -sn[j] * rhs[j] + cs[j] * rhs[j + 1];
        rhs[j] = t;
    }
    for (std::size_t ii = k; ii-- > 0;) {
        const double pivot = R[ii][ii];
        if (!(std::abs(pivot) > 0.0) || !std::isfinite(pivot)) return fresh();
        double s = rhs[ii];
        for (std::size_t jj = ii + 1; jj < k; ++jj) s -= R[ii][jj] * y[jj];
        y[ii] = s / pivot;
    }
    // residual b - V_k+1 Hbar y and the combination V_k y, on the device
    DeviceVec dr(ctx, n), dc(ctx, n);
    B.scal(1.0, B.vec(0), dr);
    B.scal(y[0], B.vec(1), dc);  // k >= 1 for a usable space
    for (std::size_t i = 0; i <= k; ++i) {
        double h = 0.0;
        for (std::size_t j = 0; j < k; ++j) h += space.Hbar[i][j] * y[j];
        B.axpy(-h, B.vec(i + 1), dr);
        if (i >= 1 && i < k) B.axpy(y[i], B.vec(i + 1), dc);
    }
    DVec residual, comb;
    dr.download(residual);
    dc.download(comb);
    precond(comb, x);
    auto out = gmres(apply, precond, b, x, rtol, restart, max_iters, false, save, &residual, ctx);
    ++out.precs;
    return out;
}

}  // namespace numa::hip
