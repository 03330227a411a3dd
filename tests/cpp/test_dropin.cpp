// C++20 drop-in check of the MI355X backend: the reference's own test cases
// for the solver path (tests/test_operators.cpp:33-297 and
// tests/test_chandrasekhar_op.cpp:60-131 of kapumota/chebyshev-nonlinear-lab)
// written against numa::hip instead of numa::op.  Exit status 0 = all checks
// passed.  Needs a HIP device.
#include <cmath>
#include <cstdio>
#include <numa/hip.hpp>

using numa::op::Counted;
using numa::op::DVec;
using numa::op::NonlinearOperator;
using numa::op::SecondOrderOperator;
using namespace numa::hip;

static_assert(NonlinearOperator<Bratu3D>);
static_assert(SecondOrderOperator<Bratu3D>);
static_assert(SecondOrderOperator<Counted<Bratu3D>>);
static_assert(SecondOrderOperator<ChandrasekharOp>);

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

// -Lap_h on n^3 nodes, zero Dirichlet, on the host (test_operators.cpp:160-184)
void apply_poisson(std::size_t n, const DVec& v, DVec& out) {
    const double h2inv = (n + 1.0) * (n + 1.0);
    out.assign(n * n * n, 0.0);
    for (std::size_t i = 0; i < n; ++i)
        for (std::size_t j = 0; j < n; ++j)
            for (std::size_t k = 0; k < n; ++k) {
                const std::size_t id = (i * n + j) * n + k;
                double s = 6.0 * v[id];
                if (i > 0) s -= v[id - n * n];
                if (i + 1 < n) s -= v[id + n * n];
                if (j > 0) s -= v[id - n];
                if (j + 1 < n) s -= v[id + n];
                if (k > 0) s -= v[id - 1];
                if (k + 1 < n) s -= v[id + 1];
                out[id] = h2inv * s;
            }
}

template <class F>
bool throws_invalid(F&& f) {
    try {
        f();
    } catch (const std::invalid_argument&) {
        return true;
    }
    return false;
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
    {  // GMRES reports a breakdown on a null operator (test_operators.cpp:141-148)
        const DVec b = {1.0, -2.0, 3.0};
        DVec x(b.size(), 0.0);
        auto apply = [](const DVec& v, DVec& out) { out.assign(v.size(), 0.0); };
        const auto r = gmres(apply, IdentityPrecond{}, b, x, 1e-10, 5, 20);
        CHECK(!r.converged && r.breakdown);
    }
    {  // restart = 0 is rejected without looping (test_operators.cpp:150-157)
        const DVec b = {1.0, 2.0};
        DVec x(b.size(), 0.0);
        auto apply = [](const DVec& v, DVec& out) { out = v; };
        const auto r = gmres(apply, IdentityPrecond{}, b, x, 1e-10, 0, 20);
        CHECK(!r.converged && r.breakdown);
    }
    {  // multigrid hierarchy and one V-cycle halve the residual (:189-205)
        PoissonMG mg(15);
        CHECK(mg.levels() == 4);
        const DVec b = make_point(15 * 15 * 15, 1.0);
        DVec z, Az;
        mg(b, z);
        apply_poisson(15, z, Az);
        for (std::size_t i = 0; i < Az.size(); ++i) Az[i] -= b[i];
        CHECK(norm2(Az) < 0.5 * norm2(b));
    }
    {  // GMRES with multigrid: iterations bounded under refinement (:207-233)
        std::size_t its[2] = {0, 0};
        const std::size_t grids[2] = {15, 31};
        for (int g = 0; g < 2; ++g) {
            const std::size_t n = grids[g];
            PoissonMG mg(n);
            const DVec b = make_point(n * n * n, 1.0);
            DVec x(n * n * n, 0.0);
            auto apply = [&](const DVec& v, DVec& out) { apply_poisson(n, v, out); };
            const auto r = gmres(apply, mg, b, x, 1e-8, 50, 200);
            CHECK(r.converged);
            its[g] = r.iters;
        }
        CHECK(its[1] <= its[0] + 3);
    }
    {  // recycled second solve without hidden products (:235-275)
        Bratu3D base(15, 1.0);
        Counted<Bratu3D> counted(base);
        const DVec x(base.size(), 0.1);
        PoissonMG mg(15);
        auto apply = [&](const DVec& v, DVec& out) { counted.jvp(x, v, out); };
        const DVec b1 = make_point(base.size(), 1.0);
        DVec b2(base.size());
        for (std::size_t i = 0; i < b2.size(); ++i) b2[i] = b1[i] * b1[i];
        DVec z(base.size(), 0.0);
        KrylovSpace space;
        const auto first = gmres(apply, mg, b1, z, 1e-8, 50, 100, true, &space);
        CHECK(first.converged && space.k > 0 && space.k <= first.iters);
        counted.reset();
        DVec u;
        const auto second = gmres_recycled(apply, mg, b2, u, 1e-8, 50, 100, space);
        CHECK(second.converged);
        CHECK(counted.counters().n_JVP == second.matvecs);
        DVec Ju;
        base.jvp(x, u, Ju);
        CHECK(relative_error(Ju, b2) < 2e-8);
        // the same on the device-resident Jacobian: the relation saved by the
        // device GMRES, recycled on the device
        KrylovSpace dspace;
        DVec zd(base.size(), 0.0), ud;
        const auto d1 = gmres(base, x, PoissonMGProvider(15), b1, zd, 1e-8, 50, 100, true, &dspace);
        CHECK(d1.converged && dspace.k > 0 && dspace.k <= d1.iters);
        CHECK(dspace.V.size() == dspace.k + 1 && dspace.V[0].size() == base.size());
        const auto d2 = gmres_recycled(base, x, PoissonMGProvider(15), b2, ud, 1e-8, 50, 100,
                                       dspace);
        CHECK(d2.converged && d2.matvecs <= second.matvecs + 2);
        base.jvp(x, ud, Ju);
        CHECK(relative_error(Ju, b2) < 2e-8);
        // r_init: the residual given, no product spent on it
        DVec zr(base.size(), 0.0);
        const auto d3 = gmres(base, x, PoissonMGProvider(15), b1, zr, 1e-8, 50, 100, true,
                              nullptr, &b1);
        CHECK(d3.converged && d3.matvecs == d1.matvecs);
    }
    {  // recycling Chebyshev keeps the wrapper's accounting (:277-297)
        Bratu3D base(15, 1.0);
        Counted<Bratu3D> counted(base);
        KrylovParams params;
        params.recycle = true;
        const auto r = chebyshev_inexact(counted, DVec(base.size(), 0.0), params,
                                         PoissonMGProvider(15));
        CHECK(r.converged);
        CHECK(r.n_JVP == counted.counters().n_JVP);
        CHECK(r.n_F == counted.counters().n_F);
        CHECK(r.n_D2 == counted.counters().n_D2);
        CHECK(r.n_prec > 0);
    }
    {  // providers bind at a point: make(x) (bratu3d.hpp:104-120, poisson_mg.hpp:303-320)
        Bratu3D f(8, 6.0);
        const DVec x = make_point(f.size(), 0.2), r = make_point(f.size(), 1.0);
        DVec z, d;
        BratuJacobi{f}.make(x)(r, z);
        f.jacobi_diag(x, d);
        bool same = true;
        for (std::size_t i = 0; i < z.size(); ++i) same = same && z[i] == r[i] / d[i];
        CHECK(same);
        const PoissonMGProvider pc(8);
        const auto M = pc.make(x);
        DVec z1, z2;
        M(r, z1);
        PoissonMG(8)(r, z2);
        CHECK(z1 == z2);
        DVec zi;
        NoPrecond{}.make(x)(r, zi);
        CHECK(zi == r);
    }
    {  // ChandrasekharOp rejects bad arguments (test_chandrasekhar_op.cpp:60-73)
        CHECK(throws_invalid([] { ChandrasekharOp(0, 0.5); }));
        CHECK(throws_invalid([] { ChandrasekharOp(10, -0.1); }));
        CHECK(throws_invalid([] { ChandrasekharOp(10, 1.1); }));
        ChandrasekharOp problem(10, 0.5);
        DVec out;
        CHECK(throws_invalid([&] { problem.eval(DVec(9, 1.0), out); }));
    }
    {  // J v and F'' consistent with differences of F (test_chandrasekhar_op.cpp:75-99)
        ChandrasekharOp problem(24, 0.8);
        DVec x(24), direction(24);
        for (std::size_t i = 0; i < 24; ++i) {
            x[i] = 1.0 + 0.2 * std::sin(0.37 * double(i) + 0.5);
            direction[i] = 0.5 * std::cos(0.23 * double(i) + 0.1);
        }
        DVec analytical, finite_difference;
        problem.jvp(x, direction, analytical);
        jvp_fd(problem, x, direction, finite_difference);
        CHECK(relative_error(finite_difference, analytical) < 2e-6);
        problem.d2(x, direction, analytical);
        d2_fd(problem, x, direction, finite_difference);
        CHECK(relative_error(finite_difference, analytical) < 2e-4);
    }
    {  // the thesis' N = 8, c = 1/2 case (test_chandrasekhar_op.cpp:101-131)
        ChandrasekharOp problem(8, 0.5);
        const DVec expected{1.021719731462, 1.073186381734, 1.125724893657, 1.169753312169,
                            1.203071751305, 1.226490874633, 1.241524600594, 1.249448516693};
        DVec residual;
        problem.eval(expected, residual);
        CHECK(norm2(residual) < 2e-10);
        const DVec x0(8, 1.0);
        const auto nk = newton_krylov(problem, x0, KrylovParams{});
        const auto cheb = chebyshev_inexact(problem, x0, KrylovParams{});
        CHECK(nk.converged && cheb.converged);
        CHECK(cheb.n_D2 >= 1);
        CHECK(relative_error(nk.x, expected) < 2e-7);
        CHECK(relative_error(cheb.x, expected) < 2e-7);
    }
    std::printf("%s (%d failures)\n", failures == 0 ? "dropin ok" : "dropin FAILED", failures);
    return failures == 0 ? 0 : 1;
}
