"""Campaign reporting of the MI355X solver in the reference's raw CSV schema.

The reference records every solve as one line of a 43-column CSV
(instrument/csv.hpp:11-118, filled by apps/run_campaign.cpp:344-475); its
W_eq / selection tooling reads those files.  This module writes the same
columns in the same formats (%.17g, %.9e, ...) from a device solve: the
counters of numa_krylov_result and the per-primitive device timers of
numa_hip_solver_timers (the Timed / TimedPrecond wrappers,
operators/timing.hpp:37-108).

    row = solve_row(ctx, n=31, lam=2.97, method="nk", precond="mg", ...)
    append_csv("c1_gpu.csv", row)
"""
from __future__ import annotations

import os
import resource
import subprocess
import time

import numpy as np

import numa_hip as H

# instrument/csv.hpp:57-64
CSV_COLUMNS = ("phase,problem,family,n,N,param,instance_id,method,precond,start_kind,start_id,"
               "distance,rho_x,seed,success,nonlinear_iters,n_F,n_JVP,n_D2,n_lin,"
               "n_lin_failures,n_prec,"
               "t_setup,t_F,t_JVP,t_D2,t_prec,t_total,max_rss_kb,residual_rel,"
               "threads,repetition,rtol,max_nl,eta_init,eta_max,gamma,alpha,"
               "eta2_factor,eta2_max,restart,maxlin,commit").split(",")

# printf conversion of every column (instrument/csv.hpp:68-73)
_FMT = {
    "phase": "s", "problem": "s", "family": "s", "n": "zu", "N": "zu", "param": ".17g",
    "instance_id": "s", "method": "s", "precond": "s", "start_kind": "s", "start_id": "s",
    "distance": "s", "rho_x": ".17g", "seed": "u", "success": "d", "nonlinear_iters": "zu",
    "n_F": "zu", "n_JVP": "zu", "n_D2": "zu", "n_lin": "zu", "n_lin_failures": "zu",
    "n_prec": "zu", "t_setup": ".9e", "t_F": ".9e", "t_JVP": ".9e", "t_D2": ".9e",
    "t_prec": ".9e", "t_total": ".9e", "max_rss_kb": "ld", "residual_rel": ".9e",
    "threads": "d", "repetition": "d", "rtol": ".17g", "max_nl": "zu", "eta_init": ".17g",
    "eta_max": ".17g", "gamma": ".17g", "alpha": ".17g", "eta2_factor": ".17g",
    "eta2_max": ".17g", "restart": "zu", "maxlin": "zu", "commit": "s",
}


def csv_header() -> str:
    return ",".join(CSV_COLUMNS)


def _fmt(conv: str, v) -> str:
    if conv == "s":
        return str(v)
    if conv in ("zu", "u", "d", "ld"):
        return "%d" % int(v)
    return ("%" + conv) % float(v)


def csv_line(row: dict) -> str:
    """One CSV line, each column converted as instrument/csv.hpp:66-118 does."""
    return ",".join(_fmt(_FMT[c], row[c]) for c in CSV_COLUMNS)


def parse_line(line: str) -> dict:
    """Inverse of csv_line for the first 43 columns of a recorded line."""
    f = line.split(",")[:len(CSV_COLUMNS)]
    row = {}
    for c, v in zip(CSV_COLUMNS, f):
        conv = _FMT[c]
        row[c] = v if conv == "s" else (int(v) if conv in ("zu", "u", "d", "ld") else float(v))
    return row


def append_csv(path: str, row: dict) -> None:
    """instrument/csv.hpp:120-137: header if the file is new or empty."""
    header = not os.path.exists(path) or os.path.getsize(path) == 0
    with open(path, "a") as f:
        if header:
            f.write(csv_header() + "\n")
        f.write(csv_line(row) + "\n")


def _commit() -> str:
    try:
        here = os.path.dirname(os.path.abspath(__file__))
        return subprocess.run(["git", "-C", here, "rev-parse", "HEAD"], capture_output=True,
                              text=True, timeout=10).stdout.strip() or "unknown"
    except Exception:
        return "unknown"


_PRECOND = {"none": H.PRECOND_NONE, "jacobi": H.PRECOND_JACOBI, "mg": H.PRECOND_GMG}


def solve_row(ctx, n: int, lam: float, method: str = "nk", precond: str = "mg",
              params: "H.KrylovParams | None" = None, x0: np.ndarray | None = None,
              precision: int = H.PRECISION_FP64, mg: "H.MGConfig | None" = None,
              jv_mode: int = H.DERIV_ANALYTIC, d2_mode: int = H.DERIV_ANALYTIC, phase: str = "gpu",
              family: str = "B1", repetition: int = 1, seed: int = 0,
              commit: str | None = None) -> tuple[dict, np.ndarray]:
    """Solves Bratu3D on the device as run_campaign.cpp:344-475 solves one
    instance (method nk | cheb | rec | tf2018, precond none | jacobi | mg; by default
    the operator's analytic J v and F'' and the reference PoissonMG, as the
    campaign runs them) and returns its CSV row and the solution.  t_setup is
    the solver (provider) construction plus the per-point set-up; times are
    device times."""
    p = (H.KrylovParams.from_buffer_copy(params) if params is not None
         else H.krylov_params())
    p.recycle = 1 if method == "rec" else 0
    methods = {"nk": H.METHOD_NEWTON, "cheb": H.METHOD_CHEBYSHEV, "rec": H.METHOD_CHEBYSHEV,
               "tf2018": H.METHOD_TF2018}
    if method not in methods:
        raise ValueError(f"unknown method {method!r} (nk | cheb | rec | tf2018)")
    cfg = H.solver_config(method=methods[method],
                          precond=_PRECOND[precond], precision=precision,
                          flexible=1 if precision == H.PRECISION_MIXED else 0,
                          jv_mode=jv_mode, d2_mode=d2_mode,
                          mg=mg if mg is not None else H.mg_reference())
    t_start = time.perf_counter()
    s = H.BratuSolver(ctx, n, lam, cfg, restart=max(1, p.gmres_restart))
    s.set_timing(True)
    static_setup = time.perf_counter() - t_start
    x = np.zeros(n ** 3) if x0 is None else np.array(x0, np.float64)
    r = s.solve_host(p, x)
    total = time.perf_counter() - t_start
    t = s.timers()
    s.close()
    row = {
        "phase": phase, "problem": "bratu", "family": family, "n": n, "N": n ** 3, "param": lam,
        "instance_id": "lambda=%f" % lam, "method": method, "precond": precond,
        "start_kind": "natural" if x0 is None else "given",
        "start_id": "natural" if x0 is None else "given", "distance": "natural",
        "rho_x": 0.0, "seed": seed, "success": int(r.converged),
        "nonlinear_iters": r.iters, "n_F": r.n_F, "n_JVP": r.n_JVP, "n_D2": r.n_D2,
        "n_lin": r.n_lin_iters, "n_lin_failures": r.n_lin_failures, "n_prec": r.n_prec,
        "t_setup": static_setup + t.t_setup_dynamic, "t_F": t.t_F, "t_JVP": t.t_JVP,
        "t_D2": t.t_D2, "t_prec": t.t_prec, "t_total": total,
        "max_rss_kb": resource.getrusage(resource.RUSAGE_SELF).ru_maxrss,
        "residual_rel": r.res / r.res0 if r.res0 > 0 else r.res, "threads": 1,
        "repetition": repetition, "rtol": p.rtol_nl, "max_nl": p.max_nl,
        "eta_init": p.eta_init, "eta_max": p.eta_max, "gamma": p.gamma, "alpha": p.alpha,
        "eta2_factor": p.eta2_factor, "eta2_max": p.eta2_max, "restart": p.gmres_restart,
        "maxlin": p.gmres_max_iters, "commit": commit if commit is not None else _commit(),
    }
    return row, x
