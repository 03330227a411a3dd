"""Per-primitive timers (operators/timing.hpp:16-108 on the device): timing
changes neither the iterates nor the counters; F, J v and the preconditioner
accumulate time, F'' only in the Chebyshev method; the timers add up to no
more than the solve's wall time; reset zeroes them; timing runs eagerly."""
import numpy as np
import pytest

import numa_hip as H

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", [H.METHOD_NEWTON, H.METHOD_CHEBYSHEV])
@pytest.mark.parametrize("precision", [H.PRECISION_FP64, H.PRECISION_MIXED])
def test_timers(ctx, method, precision):
    n, lam = 32, 6.0
    mixed = precision == H.PRECISION_MIXED
    cfg = H.solver_config(method=method, precond=H.PRECOND_GMG, precision=precision,
                          flexible=1 if mixed else 0, mg=H.mg_cheb(levels=4, fuse=H.FUSE_AUTO))
    p = H.krylov_params(gmres_restart=30)
    plain = H.BratuSolver(ctx, n, lam, cfg, restart=30)
    x0 = H.uniform_host(5, n ** 3, 0.1)
    xa = x0.copy()
    ra = plain.solve_host(p, xa)
    t0 = plain.timers()
    assert (t0.t_F, t0.t_JVP, t0.t_D2, t0.t_prec, t0.t_setup_dynamic) == (0, 0, 0, 0, 0)
    s = H.BratuSolver(ctx, n, lam, cfg, restart=30)
    s.set_timing(True)
    s.set_graphs(True)  # timing runs the steps eagerly
    xb = x0.copy()
    rb = s.solve_host(p, xb)
    np.testing.assert_array_equal(xa, xb)
    assert (ra.iters, ra.n_F, ra.n_JVP, ra.n_D2, ra.n_prec, ra.n_lin_iters) == \
        (rb.iters, rb.n_F, rb.n_JVP, rb.n_D2, rb.n_prec, rb.n_lin_iters)
    t = s.timers()
    assert t.t_F > 0 and t.t_JVP > 0 and t.t_prec > 0 and t.t_setup_dynamic > 0
    assert (t.t_D2 > 0) == (method == H.METHOD_CHEBYSHEV)
    assert t.t_F + t.t_JVP + t.t_D2 + t.t_prec + t.t_setup_dynamic <= rb.t_total
    assert s.graphs_cached() == 0
    xc = x0.copy()
    s.solve_host(p, xc)
    t2 = s.timers(reset=True)
    assert t2.t_F > t.t_F and t2.t_prec > t.t_prec  # accumulated over solves
    t3 = s.timers()
    assert (t3.t_F, t3.t_JVP, t3.t_prec) == (0, 0, 0)
    plain.close()
    s.close()


def test_campaign_row_reproduces_recorded_counters(ctx):
    """A device solve reported in the reference's CSV schema: the counters of
    the first recorded n = 31 calibration row (nk, Jacobi) are reproduced and
    the line parses back to the same values."""
    import json
    import os
    import campaign
    from conftest import GOLDEN
    cal = json.load(open(os.path.join(GOLDEN, "reference_data.json")))["calibration_n31"]
    rec = next(r for r in cal if r["method"] == "nk" and r["precond"] == "jacobi")
    p = H.krylov_params(**rec["params"])
    row, _ = campaign.solve_row(ctx, rec["n"], rec["lambda"], method="nk", precond="jacobi",
                                params=p, phase="calibration", commit="test")
    e = rec["expect"]
    assert (row["nonlinear_iters"], row["n_F"], row["n_JVP"], row["n_D2"], row["n_lin"],
            row["n_prec"]) == (e["iters"], e["n_F"], e["n_JVP"], e["n_D2"], e["n_lin_iters"],
                               e["n_prec"])
    assert row["t_F"] > 0 and row["t_JVP"] > 0 and row["t_prec"] > 0
    assert row["t_F"] + row["t_JVP"] + row["t_prec"] + row["t_D2"] <= row["t_total"]
    line = campaign.csv_line(row)
    assert campaign.csv_line(campaign.parse_line(line)) == line
    assert line.startswith("calibration,bratu,B1,31,29791,")


def test_sweep_events_time_the_finest_sweeps_without_changing_the_solve(ctx):
    """Events around the finest-level smoothing sweeps that read the iterate
    (the benchmark's roofline kernel, timed inside a solve): one per such
    launch, positive durations, and the solve's iterates unchanged."""
    n, lam = 64, 6.0
    cfg = H.solver_config(precond=H.PRECOND_GMG, precision=H.PRECISION_MIXED, flexible=1,
                          mg=H.mg_cheb(levels=4, fuse=H.FUSE_SWEEP, coarse_iters=-1))
    p = H.krylov_params(gmres_restart=30)
    x0 = H.uniform_host(3, n ** 3, 0.1)
    a = H.BratuSolver(ctx, n, lam, cfg, restart=30)
    xa = x0.copy()
    ra = a.solve_host(p, xa)
    b = H.BratuSolver(ctx, n, lam, cfg, restart=30)
    b.set_graphs(True)  # the events switch graph replay off
    b.set_sweep_events(True)
    xb = x0.copy()
    rb = b.solve_host(p, xb)
    ms, cnt = b.sweep_time()
    np.testing.assert_array_equal(xa, xb)
    assert (ra.iters, ra.n_lin_iters, ra.n_prec) == (rb.iters, rb.n_lin_iters, rb.n_prec)
    # pre-smoothing: 2 sweeps, the first from zero; post: 2 sweeps, the last
    # also sums ||z|| when the solve needs it (FD J v) -> 2 timed per V-cycle
    assert cnt == 2 * rb.n_prec, (cnt, rb.n_prec)
    assert 0.0 < ms < 1.0
    assert b.sweep_time() == (0.0, 0)  # cleared by the read
    a.close()
    b.close()
