"""Smallest-eigenpair estimator and calibrate_b6 continuation on the device.

inverse_smallest_eigenpair (analysis/spectral.hpp:34-115) runs GMRES on the
Bratu Jacobian with the solver's preconditioner: against the CPU oracle (itself
pinned bit for bit to the reference build) the eigenvalue agrees to 1e-10 and
the counters exactly or within one GMRES step; at lambda = 0 it is the closed
form 12 (n+1)^2 sin^2(pi / (2(n+1))).  The continuation of
apps/calibrate_b6.cpp:173-298 reproduces the reference's own recorded run,
experiments/c1/fold_calibration/bratu_n31_coarse.csv: all 40 natural rows up
to lambda 9.8 (lambda and gap to their 6 printed digits, corrector counts
exactly, estimator counts within 2 per row)."""
import numpy as np
import pytest

import numa_hip as H
from conftest import GOLDEN, rel

REFDATA_PATH = GOLDEN + "/reference_data.json"

pytestmark = pytest.mark.gpu


def _cfg(mg=None):
    return H.solver_config(precond=H.PRECOND_GMG, jv_mode=H.DERIV_ANALYTIC,
                           mg=mg if mg is not None else H.mg_reference())


def _ocfg(cfg):
    import oracle as O
    return O.SolverConfig(method=cfg.method, precond=cfg.precond, precision=0, flexible=0,
                          jv_mode=cfg.jv_mode, d2_mode=cfg.d2_mode,
                          mg=O.MGConfig(**{k: getattr(cfg.mg, k) for k, _ in cfg.mg._fields_}))


@pytest.mark.parametrize("n,lam,mg", [(15, 3.0, "ref"), (31, 6.5, "ref"), (32, 5.0, "cheb")])
def test_eigen_matches_oracle(ctx, oracle, n, lam, mg):
    import oracle as O
    cfg = _cfg(H.mg_cheb(levels=4) if mg == "cheb" else None)
    x = O.sin_point(n ** 3, 0.3)
    s = H.BratuSolver(ctx, n, lam, cfg, restart=100)
    vec = ctx.empty(n ** 3)
    e = s.eigen_smallest(ctx.array(x), H.eigen_params(), vec)
    s.close()
    eo, vo = oracle.bratu_eigen_smallest(n, lam, _ocfg(cfg), x)
    assert e.converged and eo.converged
    assert abs(e.eigenvalue - eo.eigenvalue) <= 1e-10 * abs(eo.eigenvalue)
    assert abs(e.outer_iterations - eo.outer_iterations) <= 1
    assert abs(e.linear_iterations - eo.linear_iterations) <= e.outer_iterations
    assert e.residual_rel <= 1e-7
    v = vec.numpy()
    assert rel(v * np.sign(v @ vo), vo) < 1e-5


@pytest.mark.parametrize("n,mg", [(31, "ref"), (32, "cheb"), (64, "cheb")])
def test_eigen_closed_form_at_lambda0(ctx, n, mg):
    cfg = _cfg(H.mg_cheb(levels=4) if mg == "cheb" else None)
    s = H.BratuSolver(ctx, n, 0.0, cfg, restart=100)
    e = s.eigen_smallest(ctx.array(np.zeros(n ** 3)))
    s.close()
    mu = 12.0 * (n + 1) ** 2 * np.sin(np.pi / (2 * (n + 1))) ** 2
    assert e.converged and abs(e.eigenvalue - mu) <= 1e-9 * mu


def test_eigen_rejects_mixed_precision_and_small_basis(ctx):
    n = 8
    s = H.BratuSolver(ctx, n, 1.0, H.solver_config(precond=H.PRECOND_GMG,
                                                    precision=H.PRECISION_MIXED, flexible=1,
                                                    mg=H.mg_cheb(levels=2)), restart=30)
    with pytest.raises(H.NumaError):
        s.eigen_smallest(ctx.array(np.zeros(n ** 3)))
    s.close()
    s = H.BratuSolver(ctx, n, 1.0, _cfg(H.mg_cheb(levels=2)), restart=30)
    with pytest.raises(H.NumaError):
        s.eigen_smallest(ctx.array(np.zeros(n ** 3)))  # GMRES(100) needs 100 vectors
    s.close()


def test_continuation_reproduces_reference_fold_rows(ctx):
    import json
    rows = [r for r in json.load(open(REFDATA_PATH))["fold_n31_coarse"]
            if r["mode"] == "natural"]
    x = ctx.empty(31 ** 3)
    pts, status, proxy = H.continuation(ctx, 31, _cfg(), H.continuation_params(lambda_stop=9.8),
                                        x_out=x)
    assert status == 0 and len(pts) == 40
    for p, r in zip(pts, rows):
        assert abs(p.lambda_ - r["lambda"]) <= 5e-6 * max(1.0, r["lambda"]), (p.lambda_, r)
        assert abs(p.spectral_gap - r["spectral_gap"]) <= 5e-6 * r["spectral_gap"], r
        assert (p.continuation_iters, p.n_F) == (r["continuation_iters"], r["n_F"]), r
        assert abs(p.n_JVP - r["n_JVP"]) <= 2 and abs(p.n_prec - r["n_prec"]) <= 2, r
    assert pts[-1].lambda_ < proxy < 10.5
    xl = x.numpy()
    assert np.all(np.isfinite(xl)) and xl.max() > 0


def test_continuation_matches_oracle_on_an_even_grid(ctx, oracle):
    import oracle as O
    cfg = _cfg(H.mg_cheb(levels=3))
    cp = H.continuation_params(lambda_stop=3.0, step=0.5)
    pts, status, proxy = H.continuation(ctx, 24, cfg, cp)
    opts, ostatus, oproxy, _ = oracle.bratu_continuation(
        24, _ocfg(cfg), O.continuation_params(lambda_stop=3.0, step=0.5))
    assert status == ostatus == 0 and len(pts) == len(opts) == 7
    for p, q in zip(pts, opts):
        assert p.lambda_ == q.lambda_
        assert abs(p.spectral_gap - q.spectral_gap) <= 1e-9 * q.spectral_gap
        assert (p.continuation_iters, p.n_F) == (q.continuation_iters, q.n_F)
    assert (np.isnan(proxy) and np.isnan(oproxy)) or abs(proxy - oproxy) <= 1e-6 * abs(oproxy)


def test_continuation_reports_corrector_failure(ctx):
    # start past the fold: every corrector fails, the step halves below
    # min_step and the run ends with calibrate_b6's exit code 3, no points
    cp = H.continuation_params(lambda_start=20.0, lambda_stop=21.0, step=0.25, min_step=0.1)
    nk = H.corrector_params()
    nk.max_nl = 20
    pts, status, proxy = H.continuation(ctx, 7, _cfg(), cp, nk)
    assert status == 3 and pts == [] and np.isnan(proxy)


def test_continuation_rejects_bad_parameters(ctx):
    with pytest.raises(H.NumaError):
        H.continuation(ctx, 7, _cfg(), H.continuation_params(lambda_start=2.0, lambda_stop=1.0))
    with pytest.raises(H.NumaError):
        H.continuation(ctx, 7, H.solver_config(precond=H.PRECOND_GMG,
                                               precision=H.PRECISION_MIXED, flexible=1,
                                               mg=H.mg_cheb(levels=2)))
