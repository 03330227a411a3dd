// continuation.hip - natural-parameter continuation of the lower Bratu3D
// branch with a spectral stop (apps/calibrate_b6.cpp:139-298) on one device.
//
// One FP64 solver (basis of at least 100 vectors: the corrector's GMRES(100)
// and the estimator's) is created for the grid and re-targeted at every
// lambda (set_lambda, a device scalar); the previous accepted solution stays
// in HBM as the next start.  Per point: Newton-Krylov corrector, then the
// smallest eigenvalue of J(x) by inverse iteration; the accepted rows are the
// reference's CSV rows (lambda, gap, eigen residual, corrector iterations,
// n_F, n_JVP + estimator matvecs, n_prec + estimator preconditioner calls).
#include <cmath>
#include <limits>
#include <memory>
#include <vector>

#include "solver_base.hpp"

namespace numa_hip {

namespace {

// Linear fit of gap^2 against lambda over the last four points; its root if
// the slope is negative and the root lies past the last point
// (calibrate_b6.cpp:139-171).
double fold_proxy(const std::vector<numa_continuation_point>& pts) {
    const double nan = std::numeric_limits<double>::quiet_NaN();
    if (pts.size() < 4) return nan;
    const size_t first = pts.size() - 4;
    double sx = 0.0, sy = 0.0, sxx = 0.0, sxy = 0.0;
    for (size_t i = first; i < pts.size(); ++i) {
        const double x = pts[i].lambda, y = pts[i].spectral_gap * pts[i].spectral_gap;
        sx += x;
        sy += y;
        sxx += x * x;
        sxy += x * y;
    }
    constexpr double count = 4.0;
    const double den = count * sxx - sx * sx;
    if (std::abs(den) < 1e-14) return nan;
    const double slope = (count * sxy - sx * sy) / den;
    const double intercept = (sy - slope * sx) / count;
    if (!(slope < 0.0)) return nan;
    const double root = -intercept / slope;
    if (!std::isfinite(root) || root <= pts.back().lambda) return nan;
    return root;
}

}  // namespace

}  // namespace numa_hip

using numa_hip::require;

extern "C" {

int numa_hip_continuation(numa_hip_ctx* ctx, int64_t n, const numa_solver_config* cfg,
                          const numa_continuation_params* cp, const numa_krylov_params* nk,
                          numa_continuation_point* points, int64_t max_points, int64_t* count,
                          int64_t* status, double* fold_proxy, double* x_out_dev) {
    NUMA_API_BEGIN
    require(ctx && ctx->impl && cfg && cp && nk && count && status && fold_proxy,
            "null argument");
    require(max_points >= 0 && (max_points == 0 || points), "bad point buffer");
    require(n >= 1 && n <= 1024, "grid size must be in [1, 1024]");
    require(cfg->precision == NUMA_PRECISION_FP64, "continuation runs an FP64 solver");
    // calibrate_b6.cpp:102-109
    require(cp->lambda_start >= 0.0 && cp->lambda_stop > cp->lambda_start && cp->step > 0.0 &&
                cp->min_step > 0.0 && cp->min_step <= cp->step && cp->spectral_stop > 0.0 &&
                cp->eigen_rtol > 0.0 && cp->linear_rtol > 0.0,
            "invalid continuation parameters");
    numa_hip::Ctx& c = *ctx->impl;
    NUMA_HIP_CHECK(hipSetDevice(c.device));
    numa_solver_config sc = *cfg;
    sc.method = NUMA_METHOD_NEWTON;
    const int64_t basis = std::max<int64_t>(100, nk->gmres_restart);
    std::unique_ptr<numa_hip::SolverBase> s(
        numa_hip::make_bratu_solver(c, n, cp->lambda_start, sc, basis));
    const int64_t N = n * n * n;
    double *prev = nullptr, *x = nullptr;
    NUMA_HIP_CHECK(hipMalloc(&prev, (size_t)N * sizeof(double)));
    struct Free {
        double* p;
        ~Free() { (void)hipFree(p); }
    } free_prev{prev};
    NUMA_HIP_CHECK(hipMalloc(&x, (size_t)N * sizeof(double)));
    Free free_x{x};
    numa_hip::launch_fill<double>(c, N, 0.0, prev);

    const numa_eigen_params ep{cp->eigen_rtol, cp->linear_rtol, 30, 100, 2000};
    std::vector<numa_continuation_point> acc;
    double lam = cp->lambda_start, step = cp->step;
    *status = 0;
    while (lam <= cp->lambda_stop + 1e-14) {
        s->set_lambda(lam);
        numa_hip::launch_convert<double, double>(c, N, prev, x);
        numa_krylov_result sol{};
        s->solve(*nk, x, sol);
        if (!sol.converged) {
            step *= 0.5;
            if (step < cp->min_step) {
                *status = 3;
                break;
            }
            lam = acc.empty() ? cp->lambda_start : acc.back().lambda + step;
            continue;
        }
        const numa_eigen_result eig = s->eigen_smallest(x, ep, nullptr);
        if (!eig.converged || !std::isfinite(eig.eigenvalue) || eig.eigenvalue <= 0.0) {
            *status = 4;
            break;
        }
        acc.push_back({lam, eig.eigenvalue, eig.residual_rel, sol.iters, sol.n_F,
                       sol.n_JVP + eig.linear_matvecs,
                       sol.n_prec + eig.preconditioner_applications});
        std::swap(prev, x);
        free_prev.p = prev;
        free_x.p = x;
        if (eig.eigenvalue <= cp->spectral_stop) break;
        if (sol.iters <= 5 && step < cp->step) step = std::min(cp->step, 1.5 * step);
        lam += step;
    }
    *count = (int64_t)acc.size();
    for (size_t i = 0; i < acc.size() && (int64_t)i < max_points; ++i) points[i] = acc[i];
    *fold_proxy = numa_hip::fold_proxy(acc);
    if (x_out_dev) numa_hip::launch_convert<double, double>(c, N, prev, x_out_dev);
    c.sync();
    NUMA_API_END
}

}  // extern "C"
