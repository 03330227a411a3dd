"""CPU pins of the oracle (oracle/numa_oracle.cpp) — no GPU.

The oracle is what every GPU parity test compares against, so it is pinned
here against:
  * the reference's own header library compiled into oracle/_ref, through the
    fixtures tests/golden/reference_vectors.npz (bit-exact where the operation
    order is the same);
  * data the reference itself recorded: the C1 internal-calibration rows
    (experiments/c1/internal_calibration/internal_calibration_raw.csv), the R1
    smoke table (docs/research/R1_OPERATOR_INTERFACE.md:113-121), the
    lower-branch solution norms (experiments/c1/starts/frozen/B1_start_hashes.csv),
    the lambda = 0 spectral gaps (experiments/c1/fold_calibration), the N = 8
    Chandrasekhar vector (tests/test_chandrasekhar_op.cpp:106-115);
  * mathematics: Laplacian sine eigenpairs, F(0) = -lambda, the Chebyshev
    residual polynomial, symmetry / contraction of the V-cycle by brute force,
    dense LU against GMRES, the Chandrasekhar moment identity.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN, rel

REFDATA = json.load(open(os.path.join(GOLDEN, "reference_data.json")))
VEC = np.load(os.path.join(GOLDEN, "reference_vectors.npz"))


def inputs(n):
    N = n ** 3
    return O.sin_point(N, 0.3), O.sin_point(N, 1.0)[::-1].copy()


# ---- Bratu3D operators -------------------------------------------------------
@pytest.mark.parametrize("n", [6, 8, 16])
@pytest.mark.parametrize("lam", [1.0, 6.5])
def test_bratu_ops_match_reference_bitwise(oracle, n, lam):
    x, v = inputs(n)
    tag = f"n{n}_l{lam}"
    assert np.array_equal(oracle.bratu_eval(n, lam, x), VEC[f"eval_{tag}"])
    assert np.array_equal(oracle.bratu_jvp(n, lam, x, v), VEC[f"jvp_{tag}"])
    assert np.array_equal(oracle.bratu_d2(n, lam, x, v), VEC[f"d2_{tag}"])
    assert np.array_equal(oracle.bratu_jacobi_diag(n, lam, x), VEC[f"diag_{tag}"])
    # FD quotients: same formulas; the reference's OpenMP norm2 may order its sum
    # differently, which moves h by an ulp
    assert rel(oracle.bratu_jvp_fd(n, lam, x, v), VEC[f"jvpfd_{tag}"]) < 1e-9
    assert rel(oracle.bratu_d2_fd(n, lam, x, v), VEC[f"d2fd_{tag}"]) < 1e-7


@pytest.mark.parametrize("n", [4, 7, 8])
def test_laplacian_sine_eigenpairs(oracle, n):
    """phi_pqr = sin(p pi (i+1)/(n+1)) ... is an eigenvector of -Lap_h with
    eigenvalue 4(n+1)^2 sum sin^2(p pi / (2(n+1)))."""
    idx = np.arange(1, n + 1)
    h2inv = (n + 1.0) ** 2
    for p, q, r in ((1, 1, 1), (2, 1, 3), (n, n, n), (1, n, 2)):
        s = [np.sin(m * np.pi * idx / (n + 1)) for m in (p, q, r)]
        phi = np.einsum("i,j,k->ijk", *s).ravel()
        mu = 4 * h2inv * sum(np.sin(m * np.pi / (2 * (n + 1))) ** 2 for m in (p, q, r))
        assert rel(oracle.poisson_apply(n, phi), mu * phi) < 1e-13
        # F(u) with lambda = 0 is the same operator
        assert rel(oracle.bratu_eval(n, 0.0, phi), mu * phi) < 1e-13


def test_F_at_zero_and_lambda0_gap(oracle):
    for n in (5, 16):
        assert np.array_equal(oracle.bratu_eval(n, 6.5, np.zeros(n ** 3)), np.full(n ** 3, -6.5))
    # the fold CSVs' lambda = 0 spectral gap is the smallest Laplacian eigenvalue
    for row in REFDATA["spectral_gaps"]:
        if row["lambda"] == 0.0:
            n = row["n"]
            mu = 12 * (n + 1.0) ** 2 * np.sin(np.pi / (2 * (n + 1))) ** 2
            assert abs(mu - row["gap"]) < 5e-4


def test_jvp_fd_and_d2_fd_against_analytic(oracle):
    """test_operators.cpp:39-61: analytic vs FD at n = 6, lambda = 1."""
    n = 6
    x = O.sin_point(n ** 3, 0.3)
    assert rel(oracle.bratu_jvp_fd(n, 1.0, x, O.sin_point(n ** 3, 1.0)),
               oracle.bratu_jvp(n, 1.0, x, O.sin_point(n ** 3, 1.0))) < 1e-6
    v = O.sin_point(n ** 3, 0.5)
    assert rel(oracle.bratu_d2_fd(n, 1.0, x, v), oracle.bratu_d2(n, 1.0, x, v)) < 1e-4


def test_dot_exact_on_integers(oracle):
    a = np.arange(1, 100001, dtype=np.float64)
    assert oracle.dot(a, a) == float(sum(int(i) * int(i) for i in range(1, 100001)))


# ---- multigrid -----------------------------------------------------------------
def test_reference_mg_levels_and_vcycle(oracle):
    assert oracle.mg_levels(15) == [15, 7, 3, 1]  # test_operators.cpp:193-194
    assert oracle.mg_levels(16) == [16]           # poisson_mg.hpp:60-62
    for n in (7, 15, 16):
        z = oracle.mg_apply(n, O.sin_point(n ** 3, 1.0))
        assert rel(z, VEC[f"mg_n{n}"]) < 1e-14
    # one V-cycle halves the residual (test_operators.cpp:189-205)
    b = O.sin_point(15 ** 3, 1.0)
    z = oracle.mg_apply(15, b)
    assert np.linalg.norm(oracle.poisson_apply(15, z) - b) < 0.5 * np.linalg.norm(b)


def cheb_cfg(levels=0, **kw):
    return O.mg_config(smoother=O.SMOOTH_CHEBYSHEV, pre=2, post=2, degree=3,
                       max_levels=levels, min_coarse_n=2, coarse_iters=24, even_coarsening=1,
                       cheb_ratio=4.0, **kw)


def wj_cfg(levels=0):
    return O.mg_config(smoother=O.SMOOTH_WJACOBI, pre=2, post=2, max_levels=levels,
                       min_coarse_n=2, coarse_iters=24, even_coarsening=1, omega=6 / 7)


def test_even_hierarchy_sizes(oracle):
    assert oracle.mg_levels(256, cheb_cfg(5)) == [256, 128, 64, 32, 16]
    assert oracle.mg_levels(32, cheb_cfg(0)) == [32, 16, 8, 4, 2]
    assert oracle.mg_levels(33, cheb_cfg(0)) == [33, 16, 8, 4, 2]


@pytest.mark.parametrize("k", [1, 3, 8, 24])
def test_chebyshev_residual_polynomial(oracle, k):
    """Single-level cycle = Chebyshev semi-iteration over the exact spectrum of
    D^-1 A from x = 0; on a sine eigenvector with D^-1 A eigenvalue t the
    residual is T_k((theta - t)/delta) / T_k(theta/delta) times the vector."""
    n = 6
    idx = np.arange(1, n + 1)
    h2inv = (n + 1.0) ** 2
    lo, hi = 1 - np.cos(np.pi / (n + 1)), 1 + np.cos(np.pi / (n + 1))
    theta, delta = (hi + lo) / 2, (hi - lo) / 2
    cfg = cheb_cfg(1, )
    cfg.coarse_iters = k
    for p, q, r in ((1, 1, 1), (3, 2, 5), (6, 6, 6)):
        phi = np.einsum("i,j,k->ijk", *[np.sin(m * np.pi * idx / (n + 1)) for m in (p, q, r)]).ravel()
        t = (4 * h2inv * sum(np.sin(m * np.pi / (2 * (n + 1))) ** 2 for m in (p, q, r))) / (6 * h2inv)
        z = oracle.mg_apply(n, phi, cfg)
        res = phi - oracle.poisson_apply(n, z)
        Tk = np.polynomial.chebyshev.Chebyshev.basis(k)
        expect = Tk((theta - t) / delta) / Tk(theta / delta)
        assert rel(res, expect * phi) < 1e-10


@pytest.mark.parametrize("name", ["reference", "wjacobi", "cheb3"])
def test_vcycle_symmetric_and_contracting(oracle, name):
    """The V-cycle M is a fixed symmetric operator (restriction = P^T/8,
    symmetric smoothing); by brute force at n = 8 (even) or 7 (reference),
    ||I - M A|| in the A-norm is well below 1."""
    n = 7 if name == "reference" else 8
    cfg = {"reference": O.mg_reference(), "wjacobi": wj_cfg(), "cheb3": cheb_cfg()}[name]
    N = n ** 3
    M = np.stack([oracle.mg_apply(n, e, cfg) for e in np.eye(N)], axis=1)
    A = np.stack([oracle.poisson_apply(n, e) for e in np.eye(N)], axis=1)
    assert np.allclose(M, M.T, atol=1e-12 * np.abs(M).max())
    # error propagation E = I - M A in the A-inner product: spectral radius
    L = np.linalg.cholesky(A)
    E = np.eye(N) - M @ A
    S = L.T @ E @ np.linalg.inv(L.T)
    rho = np.max(np.abs(np.linalg.eigvals(S)))
    assert rho < {"reference": 0.2, "wjacobi": 0.35, "cheb3": 0.1}[name]


def test_mg_gmres_iterations_bounded(oracle):
    """test_operators.cpp:207-233 for the reference, and the even-n Chebyshev
    hierarchy at 16, 32, 64."""
    its = []
    for n in (15, 31):
        _, r = oracle.poisson_gmres(n, O.sin_point(n ** 3, 1.0), 1e-8, 50, 200, O.mg_reference())
        assert r.converged
        its.append(r.iters)
    assert its[1] <= its[0] + 3
    its = []
    for n in (16, 32, 64):
        _, r = oracle.poisson_gmres(n, O.sin_point(n ** 3, 1.0), 1e-8, 30, 200, cheb_cfg())
        assert r.converged
        its.append(r.iters)
    assert max(its) <= min(its) + 2


# ---- GMRES ---------------------------------------------------------------------
@pytest.mark.parametrize("n,lam,fd,pc,restart", [(8, 1.0, 0, 2, 40), (8, 1.0, 1, 1, 40),
                                                 (12, 2.0, 0, 1, 10)])
def test_gmres_matches_reference(oracle, n, lam, fd, pc, restart):
    N = n ** 3
    cfg = O.solver_config(precond=pc, jv_mode=O.DERIV_FD if fd else O.DERIV_ANALYTIC)
    x, g = oracle.bratu_gmres(n, lam, cfg, O.sin_point(N, 0.1), O.sin_point(N, 1.0), 1e-10,
                              restart, 400)
    tag = f"gmres_n{n}_fd{fd}_pc{pc}"
    c = VEC[tag + "_counts"]
    assert (g.iters, g.matvecs, g.precs, g.converged, g.breakdown) == tuple(int(v) for v in c)
    assert rel(x, VEC[tag]) < 1e-11


def test_gmres_against_dense_lu(oracle):
    n, lam = 4, 2.0
    N = n ** 3
    x0 = O.sin_point(N, 0.2)
    J = np.stack([oracle.bratu_jvp(n, lam, x0, e) for e in np.eye(N)], axis=1)
    b = O.sin_point(N, 1.0)
    cfg = O.solver_config(precond=O.PRECOND_JACOBI, jv_mode=O.DERIV_ANALYTIC)
    x, g = oracle.bratu_gmres(n, lam, cfg, x0, b, 1e-13, 64, 200)
    assert g.converged
    assert rel(x, np.linalg.solve(J, b)) < 1e-11


def test_gmres_breakdown_and_restart_zero(oracle):
    """test_operators.cpp:141-157 (null operator via lambda = 0 and v in the
    kernel is not available; use restart = 0 and a zero RHS)."""
    n = 4
    N = n ** 3
    cfg = O.solver_config(precond=O.PRECOND_NONE, jv_mode=O.DERIV_ANALYTIC)
    x, g = oracle.bratu_gmres(n, 1.0, cfg, np.zeros(N), O.sin_point(N, 1.0), 1e-10, 0, 20)
    assert g.breakdown and not g.converged
    x, g = oracle.bratu_gmres(n, 1.0, cfg, np.zeros(N), np.zeros(N), 1e-10, 10, 20)
    assert g.converged and np.all(x == 0)


def test_fgmres_equals_gmres_for_linear_preconditioner(oracle):
    n = 16
    N = n ** 3
    x0, b = O.sin_point(N, 0.1), O.sin_point(N, 1.0)
    res = []
    for flex in (0, 1):
        cfg = O.solver_config(precond=O.PRECOND_GMG, flexible=flex, mg=cheb_cfg())
        res.append(oracle.bratu_gmres(n, 3.0, cfg, x0, b, 1e-10, 30, 200))
    (x0_, g0), (x1_, g1) = res
    assert g0.iters == g1.iters and g1.precs == g0.precs - 1
    assert rel(x1_, x0_) < 1e-12


# ---- Newton / Chebyshev --------------------------------------------------------
@pytest.mark.parametrize("n,lam,method,pc,rec", [(8, 1.0, 0, 1, 0), (15, 1.0, 1, 2, 1),
                                                 (16, 6.0, 1, 1, 0)])
def test_solves_match_reference(oracle, n, lam, method, pc, rec):
    cfg = O.solver_config(method=method, precond=pc, jv_mode=O.DERIV_ANALYTIC,
                          d2_mode=O.DERIV_ANALYTIC)
    x, r = oracle.bratu_solve(n, lam, cfg, O.krylov_params(recycle=rec))
    tag = f"solve_n{n}_m{method}_pc{pc}_r{rec}"
    c = VEC[tag + "_counts"]
    assert (r.iters, r.n_F, r.n_JVP, r.n_D2, r.n_lin_iters, r.n_prec) == tuple(int(v) for v in c)
    assert rel(x, VEC[tag]) < 1e-12


CAL = REFDATA["calibration_n31"]


@pytest.mark.parametrize("row", CAL, ids=[f"{r['method']}-{r['precond']}-l{r['lambda']:.3f}-e{r['params']['eta_init']}"
                                          for r in CAL])
def test_reference_calibration_rows(oracle, row):
    """Every counter and the final relative residual of the reference's own
    recorded C1 calibration runs (analytic J v, u0 = 0)."""
    p = O.krylov_params(recycle=1 if row["method"] == "rec" else 0, **row["params"])
    cfg = O.solver_config(method=0 if row["method"] == "nk" else 1,
                          precond=O.PRECOND_JACOBI if row["precond"] == "jacobi" else O.PRECOND_GMG,
                          jv_mode=O.DERIV_ANALYTIC, d2_mode=O.DERIV_ANALYTIC)
    _, r = oracle.bratu_solve(row["n"], row["lambda"], cfg, p)
    e = row["expect"]
    got = {k: getattr(r, k) for k in ("iters", "n_F", "n_JVP", "n_D2", "n_lin_iters",
                                      "n_lin_failures", "n_prec")}
    assert got == {k: e[k] for k in got}
    # counters are exact; the final ||F|| (~1e-9 ||F0||) sits at the roundoff
    # floor of F (the reference summed norms in one thread, the oracle in fixed
    # chunks), amplified near the fold (lambda -> 9.9) where J is nearly singular
    tol = 5e-3 if row["lambda"] > 9.0 else 1e-3
    assert abs(r.res / r.res0 - e["residual_rel"]) <= tol * e["residual_rel"]


@pytest.mark.parametrize("pc", ["mg", "jacobi"])
def test_r1_smoke_table(oracle, pc):
    t = REFDATA["r1_smoke_n16"]
    for m, want in t[pc].items():
        cfg = O.solver_config(method=0 if m == "nk" else 1,
                              precond=O.PRECOND_GMG if pc == "mg" else O.PRECOND_JACOBI,
                              jv_mode=O.DERIV_ANALYTIC, d2_mode=O.DERIV_ANALYTIC)
        _, r = oracle.bratu_solve(16, 1.0, cfg, O.krylov_params(recycle=1 if m == "rec" else 0))
        got = [r.iters, r.n_F, r.n_JVP, r.n_D2, r.n_prec]
        for g, w in zip(got, want):
            if w is not None:
                assert g == w, (m, got, want)


def test_reference_solution_norms(oracle):
    """||u*(lambda)|| of the reference's lower-branch continuation solutions."""
    rows = [r for r in REFDATA["start_norms"] if r["n"] == 31 and r["lambda"] < 6]
    assert len(rows) == 3
    for r in rows:
        cfg = O.solver_config(precond=O.PRECOND_GMG, jv_mode=O.DERIV_ANALYTIC)
        x, res = oracle.bratu_solve(31, r["lambda"], cfg, O.krylov_params(rtol_nl=1e-13))
        assert res.converged
        assert abs(np.linalg.norm(x) - r["norm"]) < 1e-9 * r["norm"]


def test_constant_forcing_converges(oracle):
    cfg = O.solver_config(precond=O.PRECOND_GMG, mg=cheb_cfg(4))
    x, r = oracle.bratu_solve(32, 6.0, cfg, O.krylov_params(gmres_restart=30, eta_init=1e-3,
                                                             forcing=O.FORCING_CONSTANT))
    assert r.converged and r.res <= 1e-8 * r.res0
    assert r.n_lin_failures == 0


# ---- Chandrasekhar -------------------------------------------------------------
def test_gauss_legendre(oracle):
    for N in (8, 24):
        a, w = oracle.gauss_legendre_unit(N)
        assert rel(a, VEC[f"gl_nodes_{N}"]) < 1e-15 and rel(w, VEC[f"gl_weights_{N}"]) < 1e-15
    a, w = oracle.gauss_legendre_unit(8)
    for d in range(16):  # test_chandrasekhar_op.cpp:36-55
        assert abs(np.sum(w * a ** d) - 1 / (d + 1)) < 5e-14


def test_chandra_ops_match_reference(oracle):
    for N in (8, 24):
        x = 1.0 + 0.2 * np.sin(0.37 * np.arange(N) + 0.5)
        d = 0.5 * np.cos(0.23 * np.arange(N) + 0.1)
        assert rel(oracle.chandra_eval(N, 0.8, x), VEC[f"chandra_eval_{N}"]) < 1e-15
        assert rel(oracle.chandra_jvp(N, 0.8, x, d), VEC[f"chandra_jvp_{N}"]) < 1e-15
        assert rel(oracle.chandra_d2(N, 0.8, x, d), VEC[f"chandra_d2_{N}"]) < 1e-15


def test_chandra_golden_and_moment_identity(oracle):
    gold = np.array(REFDATA["chandra_n8_c05"])
    assert np.linalg.norm(oracle.chandra_eval(8, 0.5, gold)) < 2e-10
    cfg = O.solver_config(method=0, precond=O.PRECOND_NONE, jv_mode=O.DERIV_ANALYTIC)
    for method in (0, 1):
        cfg.method = method
        x, r = oracle.chandra_solve(8, 0.5, cfg, O.krylov_params())
        assert r.converged and rel(x, gold) < 2e-7
    # sum_i w_i x_i = (2/c)(1 - sqrt(1 - c)) for any rule with sum w = 1
    for rule, form in ((0, 0), (1, 0), (1, 1)):
        for c in (0.5, 0.9):
            N = 64
            x, r = oracle.chandra_solve(N, c, cfg, O.krylov_params(rtol_nl=1e-13), rule=rule,
                                        form=form)
            assert r.converged, (rule, form, c)
            w = oracle.gauss_legendre_unit(N)[1] if rule == 0 else np.full(N, 1.0 / N)
            assert abs(np.dot(w, x) - 2 / c * (1 - np.sqrt(1 - c))) < 1e-11


def _fd_jvp(F, x, v):
    """Forward difference of operators/operator.hpp:101-114, from F alone."""
    h = np.sqrt(np.finfo(float).eps) * (1.0 + np.linalg.norm(x)) / np.linalg.norm(v)
    return (F(x + h * v) - F(x)) / h


def _fd_d2(F, x, v):
    """Central second difference of operators/operator.hpp:118-134, from F alone."""
    h = np.finfo(float).eps ** 0.25 * (1.0 + np.linalg.norm(x)) / np.linalg.norm(v)
    return (F(x + h * v) - 2.0 * F(x) + F(x - h * v)) / (h * h)


@pytest.mark.parametrize("rule", [0, 1])
@pytest.mark.parametrize("form", [0, 1])
def test_chandra_derivatives_consistent_with_eval(oracle, rule, form):
    """Analytic J v and F''[v,v] of both forms (reference F = 1 + x Ax - x and
    the Kelley form F = x - 1/(1 - Ax)) against differences of F itself, under
    both quadrature rules: test_chandrasekhar_op.cpp:75-99 (N = 24, c = 0.8,
    x = 1 + 0.2 sin(0.37 i + 0.5), v = 0.5 cos(0.23 i + 0.1); 2e-6 / 2e-4).
    Only chandra_eval enters the differences, so a wrong sign, power or
    operand in either derivative fails here."""
    N, c = 24, 0.8
    i = np.arange(N)
    x = 1.0 + 0.2 * np.sin(0.37 * i + 0.5)
    v = 0.5 * np.cos(0.23 * i + 0.1)
    F = lambda y: oracle.chandra_eval(N, c, y, rule=rule, form=form)  # noqa: E731
    jv = oracle.chandra_jvp(N, c, x, v, rule=rule, form=form)
    d2 = oracle.chandra_d2(N, c, x, v, rule=rule, form=form)
    assert rel(_fd_jvp(F, x, v), jv) < 2e-6
    assert rel(_fd_d2(F, x, v), d2) < 2e-4
    # the second derivative is not the first: a d2 that returned J v would pass
    # neither the bound above nor this one
    assert rel(jv, d2) > 1e-2


# ---- tensor-free Chebyshev 2018 ------------------------------------------------
def _tf_input(n):
    # test_tf_chebyshev_2018.cpp:20-25
    i = np.arange(n ** 3, dtype=np.float64)
    return 0.1 + 0.03 * np.sin(0.2 * i), 0.5 * np.cos(0.17 * i)


def _rel_err(a, b):  # test_tf_chebyshev_2018.cpp:11-15
    return np.linalg.norm(a - b) / (1.0 + np.linalg.norm(b))


@pytest.mark.parametrize("n,lam", [(4, 1.0), (8, 3.0)])
@pytest.mark.parametrize("fd", [0, 1])
def test_tf2018_curvature_matches_reference(oracle, n, lam, fd):
    x, d = _tf_input(n)
    c, h = oracle.tf2018_curvature(n, lam, x, d, fd)
    tag = f"tf2018_curv_n{n}_fd{fd}"
    assert h == VEC[tag + "_h"][0]
    # (J(x+hd)d - J(x)d)/h cancels ~8 digits: the J v products agree bit for
    # bit, so the quotient does too up to the ulp of the J v values / h
    assert _rel_err(c, VEC[tag]) < 1e-12


def test_tf2018_curvature_approximates_d2(oracle):
    # test_tf_chebyshev_2018.cpp:19-33: 0 < h <= 0.5, ||C d - F''[d,d]|| small.
    # (With the finite-difference J v the quotient is noise: each J v carries
    # ~sqrt(eps) error and h ~ sqrt(eps) ||x|| / ||d||; the reference behaves
    # the same, pinned bit for bit above.)
    n, lam = 4, 1.0
    x, d = _tf_input(n)
    c, h = oracle.tf2018_curvature(n, lam, x, d, 0)
    assert 0.0 < h <= 0.5
    assert _rel_err(c, oracle.bratu_d2(n, lam, x, d)) < 5e-6


def test_tf2018_curvature_of_a_linear_operator_is_zero(oracle):
    # lambda = 0: J does not depend on x, so J(x + h d) d - J(x) d is exactly 0;
    # d = 0 takes the step 2 eps (tf_chebyshev_2018.hpp:26)
    n = 6
    x, d = _tf_input(n)
    c, h = oracle.tf2018_curvature(n, 0.0, x, d, 0)
    assert np.all(c == 0.0) and 0.0 < h <= 0.5
    c0, h0 = oracle.tf2018_curvature(n, 1.0, x, np.zeros(n ** 3), 0)
    assert h0 == 2.0 * np.finfo(float).eps and np.all(c0 == 0.0)


TF_SOLVES = [(4, 1.0, 1, 0), (8, 3.0, 1, 0), (15, 2.0, 2, 0), (8, 1.0, 1, 1), (7, 6.0, 0, 0)]


@pytest.mark.parametrize("n,lam,pc,fd", TF_SOLVES)
def test_tf2018_solves_match_reference(oracle, n, lam, pc, fd):
    cfg = O.solver_config(method=O.METHOD_TF2018, precond=pc,
                          jv_mode=O.DERIV_FD if fd else O.DERIV_ANALYTIC,
                          d2_mode=O.DERIV_ANALYTIC, mg=O.mg_reference())
    p = O.krylov_params(rtol_nl=1e-8, max_nl=20, gmres_restart=30, gmres_max_iters=300)
    x, r = oracle.bratu_solve(n, lam, cfg, p)
    tag = f"tf2018_n{n}_l{lam}_pc{pc}_fd{fd}"
    c = VEC[tag + "_counts"]
    got = (r.converged, r.iters, r.n_F, r.n_JVP, r.n_D2, r.n_lin_iters, r.n_lin_failures, r.n_prec)
    assert got == tuple(int(v) for v in c)
    assert rel(x, VEC[tag]) < 1e-12
    if n == 4:  # test_tf_chebyshev_2018.cpp:36-48
        assert r.converged and r.n_D2 == 0 and r.n_JVP > 0


# ---- smallest eigenpair and continuation (C1 fold calibration) -----------------
@pytest.mark.parametrize("n,lam", [(15, 3.0), (7, 0.0)])
def test_eigen_matches_reference(oracle, n, lam):
    cfg = O.solver_config(precond=O.PRECOND_GMG, jv_mode=O.DERIV_ANALYTIC, mg=O.mg_reference())
    e, v = oracle.bratu_eigen_smallest(n, lam, cfg, O.sin_point(n ** 3, 0.3))
    r = VEC[f"eig_n{n}_l{lam}_res"]
    assert (e.converged, e.outer_iterations, e.linear_iterations, e.linear_matvecs,
            e.preconditioner_applications, e.linear_failures) == tuple(int(t) for t in r[[0, 3, 4, 5, 6, 7]])
    assert abs(e.eigenvalue - r[1]) <= 1e-13 * abs(r[1])
    assert abs(e.residual_rel - r[2]) <= 1e-6 * r[2]
    assert rel(v, VEC[f"eig_n{n}_l{lam}"]) < 1e-9


@pytest.mark.parametrize("n", [7, 15, 31])
def test_eigen_closed_form_at_lambda0(oracle, n):
    # lambda = 0: J = (n+1)^2 A, smallest eigenvalue 12 (n+1)^2 sin^2(pi / (2(n+1)))
    # (29.585039 at n = 31, the fold CSVs' first gap)
    cfg = O.solver_config(precond=O.PRECOND_GMG, jv_mode=O.DERIV_ANALYTIC, mg=O.mg_reference())
    e, v = oracle.bratu_eigen_smallest(n, 0.0, cfg, np.zeros(n ** 3))
    mu = 12.0 * (n + 1) ** 2 * np.sin(np.pi / (2 * (n + 1))) ** 2
    assert e.converged and abs(e.eigenvalue - mu) <= 1e-9 * mu
    k = np.arange(1, n + 1) * np.pi / (n + 1)
    phi = np.einsum("i,j,k->ijk", np.sin(k), np.sin(k), np.sin(k)).ravel()
    assert abs(abs(np.dot(v, phi)) / np.linalg.norm(phi) - 1.0) < 1e-6


def test_continuation_reproduces_reference_fold_rows(oracle):
    # every natural-continuation row of fold_calibration/bratu_n31_coarse.csv
    # up to lambda 9.8 (40 points, step halving near the fold): lambda and gap
    # to their 6 printed digits, the counters exactly
    rows = [r for r in REFDATA["fold_n31_coarse"] if r["mode"] == "natural"]
    cfg = O.solver_config(precond=O.PRECOND_GMG, jv_mode=O.DERIV_ANALYTIC, mg=O.mg_reference())
    pts, status, proxy, x = oracle.bratu_continuation(31, cfg,
                                                      O.continuation_params(lambda_stop=9.8))
    assert status == 0 and len(pts) == 40
    for p, r in zip(pts, rows):
        assert abs(p.lambda_ - r["lambda"]) <= 5e-6 * max(1.0, r["lambda"])
        assert abs(p.spectral_gap - r["spectral_gap"]) <= 5e-6 * r["spectral_gap"]
        assert abs(p.eigen_residual - r["eigen_residual"]) <= 5e-3 * r["eigen_residual"]
        assert (p.continuation_iters, p.n_F, p.n_JVP, p.n_prec) == \
            (r["continuation_iters"], r["n_F"], r["n_JVP"], r["n_prec"])
    # the gap^2 fit over the last 4 points lies past the last point
    assert pts[-1].lambda_ < proxy < 10.5
    assert np.all(np.isfinite(x)) and x.max() > 0


# ---- B1 coarse start: prolong3d --------------------------------------------------
@pytest.mark.parametrize("nc", [1, 3, 5])
def test_prolong3d_is_the_tensor_product_interpolation(nc):
    # materialize_coarse_start.cpp:60-167 against the Kronecker product of the
    # 1-D interpolation matrix (a different evaluation of the same operator)
    nf = 2 * nc + 1
    P1 = np.zeros((nf, nc))
    for f in range(nf):
        if f % 2 == 1:
            P1[f, (f - 1) // 2] = 1.0
        else:
            if f // 2 >= 1:
                P1[f, f // 2 - 1] = 0.5
            if f // 2 < nc:
                P1[f, f // 2] = 0.5
    P = np.kron(np.kron(P1, P1), P1)
    c = np.random.default_rng(nc).normal(size=nc ** 3)
    fine = O.prolong3d(c, nc)
    assert rel(fine, P @ c) < 1e-15
    # injection at the odd (coarse) nodes
    F = fine.reshape(nf, nf, nf)
    assert np.array_equal(F[1::2, 1::2, 1::2].ravel(), c)



# ---- exact coarse solve (sine transform) -----------------------------------------
@pytest.mark.parametrize("n", [1, 2, 5, 16])
def test_exact_coarse_solve_is_the_inverse(oracle, n):
    # a one-level hierarchy with coarse_iters < 0 applies A^-1: A z = r to
    # rounding, and z matches a dense solve on small grids
    cfg = O.mg_config(smoother=2, max_levels=1, coarse_iters=-1, even_coarsening=1)
    r = np.random.default_rng(n).normal(size=n ** 3)
    z = oracle.mg_apply(n, r, cfg)
    assert rel(oracle.poisson_apply(n, z), r) < 1e-13
    if n <= 5:
        A = np.column_stack([oracle.poisson_apply(n, e) for e in np.eye(n ** 3)])
        assert rel(z, np.linalg.solve(A, r)) < 1e-12


# ---- final-solve safeguard -------------------------------------------------------
def test_final_solve_safeguard_makes_the_solution_path_independent(oracle):
    """Three Krylov paths (V-cycle, Jacobi, no preconditioner) at n = 24,
    lambda = 6: with the reference's forcing they stop at different points of
    ||F|| <= 1e-8 ||F0|| (solutions ~5e-9 from the solution); with the safeguard
    (tight_res 1e-5, tight_eta 1e-6) every path ends within 1e-11 of the
    solution converged to 1e-13, and the safeguard off is the reference's
    loop (the calibration-row tests above pin that)."""
    n, lam = 24, 6.0
    mg = cheb_cfg(3)
    cfgs = [O.solver_config(precond=O.PRECOND_GMG, mg=mg), O.solver_config(precond=O.PRECOND_JACOBI),
            O.solver_config(precond=O.PRECOND_NONE)]
    ref, r = oracle.bratu_solve(n, lam, cfgs[0], O.krylov_params(
        rtol_nl=1e-13, gmres_restart=30, gmres_max_iters=3000, eta_init=1e-10, forcing=O.FORCING_CONSTANT))
    assert r.converged and r.res <= 1e-13 * r.res0
    plain = [oracle.bratu_solve(n, lam, c, O.krylov_params(gmres_restart=30, gmres_max_iters=4000))
             for c in cfgs]
    safe = [oracle.bratu_solve(n, lam, c, O.krylov_params(gmres_restart=30, gmres_max_iters=4000,
                                                          tight_res=1e-5, tight_eta=1e-6))
            for c in cfgs]
    assert all(r.converged for _, r in plain + safe)
    assert max(rel(x, ref) for x, _ in plain) > 1e-9
    assert max(rel(x, ref) for x, _ in safe) < 1e-11
    for (_, a), (_, b) in zip(plain, safe):
        assert b.iters >= a.iters


# ---- failure paths --------------------------------------------------------------
FAIL = json.load(open(os.path.join(GOLDEN, "failure_paths.json")))
FAIL_KEYS = ("converged", "iters", "n_F", "n_JVP", "n_D2", "n_lin_iters", "n_lin_failures", "n_prec")


def _same_residual(got, want, converged):
    if np.isnan(want):
        return np.isnan(got)
    return abs(got - want) <= (5e-3 if converged else 1e-9) * abs(want)


@pytest.mark.parametrize("case", FAIL["solves"],
                         ids=lambda c: f"m{c['method']}-r{c['restart']}-i{c['max_iters']}-w{c['what']}-a{c['at']}")
def test_failure_paths_match_reference(oracle, case):
    """restart = 0, max_iters = 0 and a NaN in the k-th F evaluation or J v
    product, against the reference's own loops (oracle/_ref, fixture
    tests/golden/failure_paths.json): every counter and flag identical."""
    cfg = O.solver_config(method=case["method"], precond=O.PRECOND_JACOBI,
                          jv_mode=O.DERIV_ANALYTIC, d2_mode=O.DERIV_ANALYTIC)
    p = O.krylov_params(gmres_restart=case["restart"], gmres_max_iters=case["max_iters"])
    oracle.set_inject(max(case["what"], 0), case["at"] if case["what"] >= 0 else 0)
    try:
        _, r = oracle.bratu_solve(FAIL["n"], FAIL["lambda"], cfg, p)
    finally:
        oracle.set_inject(0, 0)
    want = case["result"]
    got = O.result_dict(r)
    assert {k: got[k] for k in FAIL_KEYS} == {k: want[k] for k in FAIL_KEYS}
    assert _same_residual(got["res0"], want["res0"], False)
    assert _same_residual(got["res"], want["res"], want["converged"])


CHS = json.load(open(os.path.join(GOLDEN, "chandra_solves.json")))


@pytest.mark.parametrize("case", CHS, ids=lambda c: f"n{c['n']}-c{c['c']}-m{c['method']}")
def test_chandra_solves_match_reference(oracle, case):
    """newton_krylov / chebyshev_inexact on the reference's ChandrasekharOp
    (fixture from the reference build): every counter identical, solution to
    1e-12."""
    cfg = O.solver_config(method=case["method"], precond=O.PRECOND_NONE,
                          jv_mode=O.DERIV_ANALYTIC, d2_mode=O.DERIV_ANALYTIC)
    x, r = oracle.chandra_solve(case["n"], case["c"], cfg, O.krylov_params(rtol_nl=1e-10))
    got = O.result_dict(r)
    want = case["result"]
    assert {k: got[k] for k in FAIL_KEYS} == {k: want[k] for k in FAIL_KEYS}
    assert rel(x, case["x"]) < 1e-12
