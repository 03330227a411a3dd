"""Failure paths of the device solver against the reference's own loops
(fixture tests/golden/failure_paths.json, written from oracle/_ref): GMRES with
restart = 0 (breakdown), max_iters = 0, a zero right-hand side, and Newton /
inexact-Chebyshev solves with restart = 0, max_iters = 0 or a NaN forced from
the host into the k-th F evaluation or J v product
(numa_hip_solver_inject_nan).  Every counter and flag must be the
reference's."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN
import numa_hip as H

pytestmark = pytest.mark.gpu

FAIL = json.load(open(os.path.join(GOLDEN, "failure_paths.json")))
KEYS = ("converged", "iters", "n_F", "n_JVP", "n_D2", "n_lin_iters", "n_lin_failures", "n_prec")


def _cfg(method):
    return H.solver_config(method=method, precond=H.PRECOND_JACOBI, jv_mode=H.DERIV_ANALYTIC,
                           d2_mode=H.DERIV_ANALYTIC)


def _same_residual(got, want, converged):
    if np.isnan(want):
        return np.isnan(got)
    return abs(got - want) <= (5e-3 if converged else 1e-9) * abs(want)


@pytest.mark.parametrize("case", FAIL["solves"],
                         ids=lambda c: f"m{c['method']}-r{c['restart']}-i{c['max_iters']}-w{c['what']}-a{c['at']}")
@pytest.mark.parametrize("graphs", [False, True])
def test_solve_failure_paths(ctx, case, graphs):
    n = FAIL["n"]
    print(f"failure path {case['method']} {case['restart']} {case['max_iters']} "
          f"{case['what']} {case['at']} graphs={graphs}", flush=True)
    s = H.BratuSolver(ctx, n, FAIL["lambda"], _cfg(case["method"]), restart=30)
    if graphs:
        s.set_graphs(True)
    if case["what"] >= 0:
        s.inject_nan(case["what"], case["at"])
    p = H.krylov_params(gmres_restart=case["restart"], gmres_max_iters=case["max_iters"])
    x = np.zeros(n ** 3)
    r = s.solve_host(p, x)
    s.close()
    want = case["result"]
    got = {k: int(getattr(r, k)) for k in KEYS}
    assert got == {k: int(want[k]) for k in KEYS}
    assert _same_residual(r.res0, want["res0"], False)
    assert _same_residual(r.res, want["res"], want["converged"])


def test_injection_fires_once_and_disarms(ctx):
    n = FAIL["n"]
    s = H.BratuSolver(ctx, n, FAIL["lambda"], _cfg(0), restart=30)
    p = H.krylov_params(gmres_restart=30)
    clean = s.solve_host(p, np.zeros(n ** 3))
    s.inject_nan(H.INJECT_F, 1)
    hit = s.solve_host(p, np.zeros(n ** 3))
    assert np.isnan(hit.res0) and not hit.converged
    again = s.solve_host(p, np.zeros(n ** 3))  # fired once: the next solve is clean
    assert (again.iters, again.n_JVP, again.converged) == (clean.iters, clean.n_JVP, 1)
    s.inject_nan(H.INJECT_JV, 3)
    s.inject_nan(H.INJECT_JV, 0)  # disarmed
    assert s.solve_host(p, np.zeros(n ** 3)).converged
    with pytest.raises(H.NumaError):
        s.inject_nan(7, 1)
    s.close()


@pytest.mark.parametrize("case", FAIL["gmres"], ids=lambda c: c["case"])
def test_gmres_failure_paths(ctx, case):
    n = FAIL["n"]
    N = n ** 3
    from oracle import sin_point
    b = np.zeros(N) if case["case"] == "zero_rhs" else sin_point(N, 1.0)
    s = H.BratuSolver(ctx, n, FAIL["lambda"], _cfg(0), restart=30)
    x0, db = ctx.array(sin_point(N, 0.1)), ctx.array(b)
    z = ctx.array(np.full(N, 3.0))
    g = s.gmres(x0, db, z, 1e-10, case["restart"], case["max_iters"])
    want = case["result"]
    assert (g.iters, g.matvecs, g.precs, bool(g.converged), bool(g.breakdown)) == \
        (want["iters"], want["matvecs"], want["precs"], want["converged"], want["breakdown"])
    assert g.rel_res == want["rel_res"]
    if case["case"] == "zero_rhs":  # x = 0 (gmres.hpp:71-75)
        assert np.all(z.numpy() == 0.0)
    s.close()
