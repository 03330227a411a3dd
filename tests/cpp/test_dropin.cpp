// C++20 drop-in check of the MI355X backend: the reference's own test cases
// for the solver path (tests/test_operators.cpp:33-120 and :189-297 of
// kapumota/chebyshev-nonlinear-lab) written against numa::hip instead of
// numa::op.  Exit status 0 = all checks passed.  Needs a HIP device.
#include <cmath>
#include <cstdio>
#include <numa/hip/bratu3d.hpp>

using numa::op::Counted;
using numa::op::DVec;
using numa::op::NonlinearOperator;
using numa::op::SecondOrderOperator;
using namespace numa::hip;

static_assert(NonlinearOperator<Bratu3D>);
static_assert(SecondOrderOperator<Bratu3D>);
static_assert(SecondOrderOperator<Counted<Bratu3D>>);

namespace {

int failures = 0;
#define CHECK(cond)                                                         \
    do {                                                                    \
        if (!(cond)) {                                                      \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);     \
            ++failures;                                                     \
        }                                                                   \
    } while (0)

DVec make_point(std::size_t n, double scale) {
    DVec x(n);
    for (std::size_t i = 0; i < n; ++i) x[i] = scale * std::sin(0.37 * double(i) + 0.5);
    return x;
}

double norm2(const DVec& a) {
    double s = 0;
    for (double v : a) s += v * v;
    return std::sqrt(s);
}

double relative_error(const DVec& a, const DVec& b) {
    DVec d(a.size());
    for (std::size_t i = 0; i < a.size(); ++i) d[i] = a[i] - b[i];
    return norm2(d) / norm2(b);
}

}  // namespace

int main() {
    Context ctx(0);
    {  // GMRES on the Bratu Jacobian (test_operators.cpp:78-89)
        Bratu3D f(ctx, 8, 1.0);
        const DVec x0(f.size(), 0.0);
        const DVec b = make_point(f.size(), 1.0);
        DVec sol(f.size(), 0.0);
        const auto r = gmres(f, x0, NoPrecond{}, b, sol, 1e-10, 40, 400);
        CHECK(r.converged);
        DVec Ax;
        f.jvp(x0, sol, Ax);
        CHECK(relative_error(Ax, b) < 1e-9);
    }
    {  // Newton-Krylov and inexact Chebyshev converge (test_operators.cpp:91-110)
        Bratu3D f(ctx, 8, 1.0);
        const DVec x0(f.size(), 0.0);
        const auto r = newton_krylov(f, x0, KrylovParams{}, BratuJacobi{f});
        CHECK(r.converged);
        DVec F;
        f.eval(r.x, F);
        CHECK(norm2(F) <= 1e-8 * r.res0 * 1.0001);
        const auto c = chebyshev_inexact(f, x0, KrylovParams{}, BratuJacobi{f});
        CHECK(c.converged);
        CHECK(c.n_D2 >= 1);
        f.eval(c.x, F);
        CHECK(norm2(F) <= 1e-8 * c.res0 * 1.0001);
    }
    {  // Counted wrapper works on the device operator (test_operators.cpp:112-120)
        Bratu3D base(ctx, 8, 1.0);
        Counted<Bratu3D> f(base);
        DVec out;
        f.eval(DVec(base.size(), 0.0), out);
        f.jvp(DVec(base.size(), 0.0), make_point(base.size(), 1.0), out);
        CHECK(f.counters().n_F == 1 && f.counters().n_JVP == 1);
    }
    {  // Chebyshev with recycling and multigrid (test_operators.cpp:277-297)
        Bratu3D f(ctx, 15, 1.0);
        KrylovParams p;
        p.recycle = true;
        const auto r = chebyshev_inexact(f, DVec(f.size(), 0.0), p, PoissonMGProvider(15));
        CHECK(r.converged);
        CHECK(r.n_prec > 0);
    }
    {  // R1 smoke table, n = 16, lambda = 1, multigrid (R1_OPERATOR_INTERFACE.md:113-121)
        Bratu3D f(ctx, 16, 1.0);
        const DVec x0(f.size(), 0.0);
        const auto nk = newton_krylov(f, x0, KrylovParams{}, PoissonMGProvider(16));
        CHECK(nk.iters == 4 && nk.n_F == 5 && nk.n_JVP == 11 && nk.n_prec == 15);
        const auto ch = chebyshev_inexact(f, x0, KrylovParams{}, PoissonMGProvider(16));
        CHECK(ch.iters == 4 && ch.n_JVP == 20 && ch.n_D2 == 4 && ch.n_prec == 28);
    }
    {  // Timed / TimedPrecond (operators/timing.hpp:37-108): per-primitive times
        Bratu3D f(ctx, 16, 1.0);
        const DVec x0(f.size(), 0.0);
        Timers t;
        SolverOptions o;
        o.timers = &t;
        const auto nk = newton_krylov(f, x0, KrylovParams{}, PoissonMGProvider(16), o);
        CHECK(nk.converged && nk.iters == 4 && nk.n_JVP == 11 && nk.n_prec == 15);
        CHECK(t.t_F > 0.0 && t.t_JVP > 0.0 && t.t_prec > 0.0 && t.t_D2 == 0.0);
        CHECK(t.t_F + t.t_JVP + t.t_prec + t.t_setup_dynamic <= nk.t_total);
        const Timers after_nk = t;
        const auto ch = chebyshev_inexact(f, x0, KrylovParams{}, PoissonMGProvider(16), o);
        CHECK(ch.converged && ch.n_D2 == 4);
        CHECK(t.t_D2 > 0.0 && t.t_F > after_nk.t_F && t.t_prec > after_nk.t_prec);
    }
    {  // test_tf_chebyshev_2018.cpp:19-48: curvature vs F'', TF-2018 converges
        Bratu3D f(ctx, 4, 1.0);
        DVec x(f.size()), d(f.size());
        for (std::size_t i = 0; i < f.size(); ++i) {
            x[i] = 0.1 + 0.03 * std::sin(0.2 * static_cast<double>(i));
            d[i] = 0.5 * std::cos(0.17 * static_cast<double>(i));
        }
        DVec tensor_free, exact;
        double h = 0.0;
        tensor_free_curvature_2018(f, x, d, tensor_free, &h);
        f.d2(x, d, exact);
        CHECK(h > 0.0 && h <= 0.5);
        CHECK(relative_error(tensor_free, exact) < 5e-6);
        KrylovParams p;
        p.rtol_nl = 1e-8;
        p.max_nl = 20;
        p.gmres_restart = 30;
        p.gmres_max_iters = 300;
        const auto r = tf_chebyshev_2018(f, DVec(f.size(), 0.0), p, BratuJacobi{f});
        CHECK(r.converged && r.n_D2 == 0 && r.n_JVP > 0);
    }
    {  // analysis/spectral.hpp: smallest eigenvalue of J at lambda = 0 is
       // 12 (n+1)^2 sin^2(pi / (2(n+1))) (the fold CSVs' first gap, 29.585 at n = 31)
        Bratu3D f(ctx, 31, 0.0);
        const auto e = inverse_smallest_eigenpair(f, DVec(f.size(), 0.0), PoissonMGProvider(31));
        const double s = std::sin(3.14159265358979323846 / 64.0);
        CHECK(e.converged && std::abs(e.eigenvalue - 12.0 * 1024.0 * s * s) < 1e-7);
        CHECK(e.vector.size() == f.size() && e.residual_rel <= 1e-7);
    }
    {  // misuse is reported as std::invalid_argument
        bool thrown = false;
        try {
            Bratu3D f(ctx, 8, 1.0);
            newton_krylov(f, DVec(7, 0.0));
        } catch (const std::invalid_argument&) {
            thrown = true;
        }
        CHECK(thrown);
    }
    std::printf("%s (%d failures)\n", failures == 0 ? "dropin ok" : "dropin FAILED", failures);
    return failures == 0 ? 0 : 1;
}
