"""Tensor-free Chebyshev 2018 (methods/tf_chebyshev_2018.hpp) on the device
against the CPU oracle, which is pinned bit for bit to the reference's own
tensor_free_curvature_2018 / tf_chebyshev_2018 (tests/test_oracle.py).

The curvature [J(x + h d) d - J(x) d] / h cancels about 8 digits (h ~ 4e-9),
so an ulp of the device exp() in J v shows up at ~1e-8 relative in it: the
device curvature is held to 1e-6 of the oracle's, and to the reference test's
5e-6 of the analytic F''[d, d] on small grids (on larger ones J v grows as
(n+1)^2 and so does the quotient's rounding, for the oracle too).  FP64
solves must reach the oracle's outer count, F count and convergence flag,
with the solution within 1e-9 relative and inner GMRES counts within 2
(preconditioned) or 25% (unpreconditioned: inner tolerances reach 2 eps)."""
import numpy as np
import pytest

import numa_hip as H
from conftest import rel

pytestmark = pytest.mark.gpu


def _tf_input(n):
    # test_tf_chebyshev_2018.cpp:20-25
    i = np.arange(n ** 3, dtype=np.float64)
    return 0.1 + 0.03 * np.sin(0.2 * i), 0.5 * np.cos(0.17 * i)


def _rel_err(a, b):  # test_tf_chebyshev_2018.cpp:11-15
    return np.linalg.norm(a - b) / (1.0 + np.linalg.norm(b))


def _curvature(ctx, n, lam, x, d, jv_mode=H.DERIV_ANALYTIC):
    cfg = H.solver_config(method=H.METHOD_TF2018, precond=H.PRECOND_NONE, jv_mode=jv_mode)
    s = H.BratuSolver(ctx, n, lam, cfg, restart=1)
    out = ctx.empty(n ** 3)
    h = s.tf2018_curvature(ctx.array(x), ctx.array(d), out)
    c = out.numpy()
    s.close()
    return c, h


@pytest.mark.parametrize("n,lam", [(4, 1.0), (8, 3.0), (33, 6.0)])
def test_curvature_matches_oracle_and_d2(ctx, oracle, n, lam):
    x, d = _tf_input(n)
    c, h = _curvature(ctx, n, lam, x, d)
    co, ho = oracle.tf2018_curvature(n, lam, x, d, 0)
    assert 0.0 < h <= 0.5 and abs(h - ho) <= 1e-13 * ho
    assert _rel_err(c, co) < 1e-6
    d2 = oracle.bratu_d2(n, lam, x, d)
    if n <= 8:  # the reference test's bar (n = 4)
        assert _rel_err(c, d2) < 5e-6
    else:  # J v ~ (n+1)^2: the quotient's rounding grows; no worse than the oracle
        assert _rel_err(c, d2) < 2.0 * _rel_err(co, d2) + 1e-6


def test_curvature_special_cases(ctx, oracle):
    n = 6
    x, d = _tf_input(n)
    # linear operator (lambda = 0): J(x + h d) d == J(x) d bit for bit
    c, h = _curvature(ctx, n, 0.0, x, d)
    assert np.all(c == 0.0) and 0.0 < h <= 0.5
    # d = 0: h = 2 eps (tf_chebyshev_2018.hpp:26), curvature 0
    c0, h0 = _curvature(ctx, n, 1.0, x, np.zeros(n ** 3))
    assert h0 == 2.0 * np.finfo(float).eps and np.all(c0 == 0.0)
    # finite-difference J v: same step h as the oracle, finite values
    cf, hf = _curvature(ctx, n, 1.0, x, d, H.DERIV_FD)
    _, hfo = oracle.tf2018_curvature(n, 1.0, x, d, 1)
    assert abs(hf - hfo) <= 1e-13 * hfo and np.all(np.isfinite(cf))


def test_curvature_rejects_aliasing(ctx):
    n = 4
    x, d = _tf_input(n)
    cfg = H.solver_config(method=H.METHOD_TF2018, precond=H.PRECOND_NONE,
                          jv_mode=H.DERIV_ANALYTIC)
    s = H.BratuSolver(ctx, n, 1.0, cfg, restart=1)
    dx, dd = ctx.array(x), ctx.array(d)
    with pytest.raises(H.NumaError):
        s.tf2018_curvature(dx, dd, dd)
    s.close()


TF_SOLVES = [  # (n, lambda, precond, mg, gmres restart)
    (4, 1.0, H.PRECOND_JACOBI, None, 30),  # test_tf_chebyshev_2018.cpp:36-48
    (8, 3.0, H.PRECOND_JACOBI, None, 30),
    (15, 2.0, H.PRECOND_GMG, "ref", 30),
    (7, 6.0, H.PRECOND_NONE, None, 200),
    (32, 6.0, H.PRECOND_GMG, "cheb", 30),
    (7, 6.0, H.PRECOND_NONE, None, 30),
]


@pytest.mark.parametrize("case", TF_SOLVES, ids=lambda c: f"n{c[0]}-l{c[1]}-pc{c[2]}-m{c[4]}")
def test_tf2018_solve_fp64_matches_oracle(ctx, oracle, case):
    import oracle as O
    n, lam, pc, mg, m = case
    mgc = H.mg_cheb(levels=3) if mg == "cheb" else H.mg_reference()
    cfg = H.solver_config(method=H.METHOD_TF2018, precond=pc, jv_mode=H.DERIV_ANALYTIC,
                          d2_mode=H.DERIV_ANALYTIC, mg=mgc)
    p = H.krylov_params(rtol_nl=1e-8, max_nl=20, gmres_restart=m, gmres_max_iters=300)
    x, r = H.tf_chebyshev_2018(ctx, n, lam, np.zeros(n ** 3), p, cfg)
    ocfg = O.SolverConfig(method=cfg.method, precond=cfg.precond, precision=0, flexible=0,
                          jv_mode=cfg.jv_mode, d2_mode=cfg.d2_mode,
                          mg=O.MGConfig(**{k: getattr(mgc, k) for k, _ in mgc._fields_}))
    op = O.KrylovParams(**{k: getattr(p, k) for k, _ in p._fields_})
    xo, ro = oracle.bratu_solve(n, lam, ocfg, op)
    assert ro.converged and r.converged
    assert (r.iters, r.n_F, r.n_D2, r.n_lin_failures) == (ro.iters, ro.n_F, 0, 0)
    # inner tolerance max(2 eps, min(1e-8, ||F||^2)) (tf_chebyshev_2018.hpp:
    # 74-77): the last inner solves run to ~1e-16, where GMRES's history
    # follows each implementation's rounding; preconditioned solves reach it
    # in a few steps
    tol = 2 if pc != H.PRECOND_NONE else max(2, int(0.25 * ro.n_lin_iters))
    assert abs(r.n_lin_iters - ro.n_lin_iters) <= tol
    assert abs(r.n_JVP - ro.n_JVP) <= tol
    assert r.res <= 1e-8 * r.res0
    assert rel(x, xo) < 1e-9


def test_tf2018_counts_follow_the_method(ctx):
    # per outer step two GMRES solves and the 2 J v of the curvature; one F
    # per step plus the first; no F'' (tf_chebyshev_2018.hpp:86-126)
    n, lam = 16, 4.0
    cfg = H.solver_config(method=H.METHOD_TF2018, precond=H.PRECOND_GMG,
                          jv_mode=H.DERIV_ANALYTIC, mg=H.mg_cheb(levels=3))
    p = H.krylov_params(rtol_nl=1e-8, max_nl=20, gmres_restart=30, gmres_max_iters=300)
    _, r = H.tf_chebyshev_2018(ctx, n, lam, np.zeros(n ** 3), p, cfg)
    assert r.converged and r.n_F == r.iters + 1 and r.n_D2 == 0
    assert r.n_JVP >= r.n_lin_iters + 2 * r.iters and r.n_prec >= r.n_lin_iters


def test_tf2018_mixed_converges(ctx):
    n, lam = 32, 5.0
    cfg = H.solver_config(method=H.METHOD_TF2018, precond=H.PRECOND_GMG,
                          precision=H.PRECISION_MIXED, flexible=1, jv_mode=H.DERIV_ANALYTIC,
                          mg=H.mg_cheb(levels=3))
    p = H.krylov_params(rtol_nl=1e-8, max_nl=20, gmres_restart=30, gmres_max_iters=300)
    x, r = H.tf_chebyshev_2018(ctx, n, lam, np.zeros(n ** 3), p, cfg)
    assert r.iters >= 1 and np.all(np.isfinite(x))
    assert r.res < r.res0
