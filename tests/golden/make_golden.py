"""Writes the golden fixtures under tests/golden/.

Run in the build container, where the reference sources exist:
    python tests/golden/make_golden.py

Two kinds of fixtures:
  * reference-held data: numbers copied from the reference's own data files
    (experiments/c1/internal_calibration/internal_calibration_raw.csv,
    experiments/c1/starts/frozen/B1_start_hashes.csv,
    experiments/c1/fold_calibration/bratu_n*_coarse.csv (spectral gaps, and every
    row of the n = 31 continuation),
    docs/research/R1_OPERATOR_INTERFACE.md:113-121,
    tests/test_chandrasekhar_op.cpp:106-115);
  * vectors computed by the reference's own header library compiled into
    oracle/_ref (oracle/Makefile) on deterministic inputs.
Only data is written; no reference source is copied.
"""
import csv
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

REF = "/root/reference"


def calibration_rows():
    """NK / Chebyshev / Chebyshev+recycling rows of the C1 internal
    calibration at n = 31 from the natural start (u0 = 0)."""
    path = os.path.join(REF, "experiments/c1/internal_calibration/internal_calibration_raw.csv")
    out, seen = [], set()
    for r in csv.DictReader(open(path)):
        if r["family"] != "B1" or r["start_kind"] != "natural" or r["n"] != "31":
            continue
        if r["method"] not in ("nk", "cheb", "rec") or r["precond"] not in ("jacobi", "mg"):
            continue
        key = (r["param"], r["method"], r["precond"], r["eta_init"], r["restart"], r["maxlin"])
        if key in seen:
            continue
        seen.add(key)
        out.append({
            "source": "experiments/c1/internal_calibration/internal_calibration_raw.csv",
            "n": int(r["n"]), "lambda": float(r["param"]), "method": r["method"],
            "precond": r["precond"],
            "params": {"rtol_nl": float(r["rtol"]), "max_nl": int(r["max_nl"]),
                       "eta_init": float(r["eta_init"]), "eta_max": float(r["eta_max"]),
                       "gamma": float(r["gamma"]), "alpha": float(r["alpha"]),
                       "eta2_factor": float(r["eta2_factor"]), "eta2_max": float(r["eta2_max"]),
                       "gmres_restart": int(r["restart"]), "gmres_max_iters": int(r["maxlin"])},
            "expect": {"iters": int(r["nonlinear_iters"]), "n_F": int(r["n_F"]),
                       "n_JVP": int(r["n_JVP"]), "n_D2": int(r["n_D2"]),
                       "n_lin_iters": int(r["n_lin"]), "n_lin_failures": int(r["n_lin_failures"]),
                       "n_prec": int(r["n_prec"]), "residual_rel": float(r["residual_rel"])}})
    return out


def csv_rows(per_group: int = 4):
    """Raw result lines of the reference's campaign CSV (instrument/csv.hpp:57-118
    format, 43 columns; the calibration file appends two more): header and a
    few lines per (problem, method, precond), kept as text."""
    path = os.path.join(REF, "experiments/c1/internal_calibration/internal_calibration_raw.csv")
    lines = open(path).read().splitlines()
    header = ",".join(lines[0].split(",")[:43])
    out, count = [], {}
    for line in lines[1:]:
        f = line.split(",")
        key = (f[1], f[7], f[8])
        if count.get(key, 0) >= per_group:
            continue
        count[key] = count.get(key, 0) + 1
        out.append(",".join(f[:43]))
    return {"source": "experiments/c1/internal_calibration/internal_calibration_raw.csv",
            "header": header, "lines": out}


def start_norms():
    """Lower-branch solution norms ||u*(lambda)|| (natural-mode continuation)."""
    path = os.path.join(REF, "experiments/c1/starts/frozen/B1_start_hashes.csv")
    out = []
    for r in csv.DictReader(open(path)):
        if r["mode"] == "natural":
            out.append({"n": int(r["n"]), "lambda": float(r["reference_lambda"]),
                        "norm": float(r["reference_norm"])})
    return out


def fold_rows(n: int = 31, kind: str = "coarse"):
    """experiments/c1/fold_calibration/bratu_n{n}_{kind}.csv: every row (the
    calibrate_b6 continuation, apps/calibrate_b6.cpp:128-137 format)."""
    path = f"{REF}/experiments/c1/fold_calibration/bratu_n{n}_{kind}.csv"
    out = []
    with open(path) as f:
        for r in csv.DictReader(f):
            out.append({"lambda": float(r["lambda"]), "spectral_gap": float(r["spectral_gap"]),
                        "eigen_residual": float(r["eigen_residual"]),
                        "continuation_iters": int(r["continuation_iters"]),
                        "n_F": int(r["n_F"]), "n_JVP": int(r["n_JVP"]),
                        "n_prec": int(r["n_prec"]), "mode": r["continuation_mode"]})
    return out


def spectral_gaps():
    out = []
    for n in (31, 63, 127):
        path = os.path.join(REF, f"experiments/c1/fold_calibration/bratu_n{n}_coarse.csv")
        for r in csv.DictReader(open(path)):
            lam = float(r["lambda"])
            if lam in (0.0, 6.5):
                out.append({"n": n, "lambda": lam, "gap": float(r["spectral_gap"])})
    return out


# (n, lambda, precond 0 none / 1 Jacobi / 2 PoissonMG, fd) of the TF-2018 solve
# fixtures; (4, 1.0, 1, 0) is test_tf_chebyshev_2018.cpp:36-48
TF_SOLVES = ((4, 1.0, 1, 0), (8, 3.0, 1, 0), (15, 2.0, 2, 0), (8, 1.0, 1, 1), (7, 6.0, 0, 0))


# Failure paths (test_operators.cpp:141-157 and the non-finite guards of
# gmres.hpp:165-169, 185-190, krylov_newton.hpp:94-104): Bratu3D n = 8,
# lambda = 6, u0 = 0, analytic derivatives, BratuJacobi, GMRES(30).
# (method, restart, max_iters, what, at): what -1 = no injection, 0 = NaN in
# the at-th F evaluation, 1 = NaN in the at-th J v product.
FAILURE_SOLVES = ((0, 0, 500, -1, 0), (0, 30, 0, -1, 0), (1, 0, 500, -1, 0),
                  (0, 30, 500, 0, 1), (0, 30, 500, 0, 2), (0, 30, 500, 0, 3),
                  (0, 30, 500, 1, 1), (0, 30, 500, 1, 7), (1, 30, 500, 0, 3),
                  (1, 30, 500, 1, 7), (1, 30, 500, 1, 30))


def failure_paths(ref):
    out = {"n": 8, "lambda": 6.0, "precond": 1, "solves": [], "gmres": []}
    for method, restart, max_iters, what, at in FAILURE_SOLVES:
        p = O.krylov_params(gmres_restart=restart, gmres_max_iters=max_iters)
        if what < 0:
            _, r = ref.bratu_solve(8, 6.0, method, 1, 0, p)
        else:
            _, r = ref.bratu_solve_inject(8, 6.0, method, 1, p, what, at)
        out["solves"].append({"method": method, "restart": restart, "max_iters": max_iters,
                              "what": what, "at": at, "result": r})
    # GMRES on J(x0) z = b: b = 0, restart = 0, max_iters = 0
    N = 8 ** 3
    for name, b, restart, max_iters in (("zero_rhs", np.zeros(N), 30, 100),
                                        ("restart0", O.sin_point(N, 1.0), 0, 100),
                                        ("maxiter0", O.sin_point(N, 1.0), 30, 0)):
        _, g = ref.bratu_gmres(8, 6.0, 0, 1, O.sin_point(N, 0.1), b, 1e-10, restart, max_iters)
        out["gmres"].append({"case": name, "restart": restart, "max_iters": max_iters,
                             "result": g})
    return out


def chandra_solves(ref):
    """newton_krylov / chebyshev_inexact on ChandrasekharOp (Gauss-Legendre,
    x0 = 1, default KrylovParams but rtol 1e-10): counters and solutions."""
    out = []
    for n in (8, 24, 64):
        for c in (0.5, 0.9, 0.9999):
            for method in (0, 1):
                x, r = ref.chandra_solve(n, c, method, O.krylov_params(rtol_nl=1e-10))
                out.append({"n": n, "c": c, "method": method, "result": r, "x": list(x)})
    return out


def main():
    with open(os.path.join(HERE, "reference_csv_rows.json"), "w") as f:
        json.dump(csv_rows(), f, indent=1)
    if "--csv-only" in sys.argv:
        return
    ref = O.Reference()
    with open(os.path.join(HERE, "failure_paths.json"), "w") as f:
        json.dump(failure_paths(ref), f, indent=1)
    with open(os.path.join(HERE, "chandra_solves.json"), "w") as f:
        json.dump(chandra_solves(ref), f, indent=1)
    if "--failure-only" in sys.argv:
        return
    data = {
        "calibration_n31": calibration_rows(),
        "start_norms": start_norms(),
        "spectral_gaps": spectral_gaps(),
        "fold_n31_coarse": fold_rows(31, "coarse"),
        # docs/research/R1_OPERATOR_INTERFACE.md:113-121 (n = 16, lambda = 1, u0 = 0)
        "r1_smoke_n16": {
            "mg": {"nk": [4, 5, 11, 0, 15], "cheb": [4, 5, 20, 4, 28], "rec": [4, 5, 18, 4, 30]},
            "jacobi": {"nk": [5, None, 74, None, 79], "cheb": [5, None, 123, None, 133],
                       "rec": [5, None, 99, None, 112]},
            "fields": ["iters", "n_F", "n_JVP", "n_D2", "n_prec"]},
        # tests/test_chandrasekhar_op.cpp:106-115 (N = 8, c = 0.5, Gauss-Legendre)
        "chandra_n8_c05": [1.021719731462, 1.073186381734, 1.125724893657, 1.169753312169,
                           1.203071751305, 1.226490874633, 1.241524600594, 1.249448516693],
    }
    with open(os.path.join(HERE, "reference_data.json"), "w") as f:
        json.dump(data, f, indent=1)

    # vectors from the reference's compiled headers on deterministic inputs
    arrays = {}
    for n in (6, 8, 15, 16):
        N = n ** 3
        x = O.sin_point(N, 0.3)
        v = O.sin_point(N, 1.0)[::-1].copy()
        for lam in (1.0, 6.5):
            tag = f"n{n}_l{lam}"
            arrays[f"eval_{tag}"] = ref.bratu("eval", n, lam, x)
            arrays[f"jvp_{tag}"] = ref.bratu("jvp", n, lam, x, v)
            arrays[f"d2_{tag}"] = ref.bratu("d2", n, lam, x, v)
            arrays[f"diag_{tag}"] = ref.bratu("jacobi_diag", n, lam, x)
            arrays[f"jvpfd_{tag}"] = ref.bratu("jvp_fd", n, lam, x, v)
            arrays[f"d2fd_{tag}"] = ref.bratu("d2_fd", n, lam, x, v)
    for n in (7, 15, 16):
        arrays[f"mg_n{n}"] = ref.mg_apply(n, O.sin_point(n ** 3, 1.0))
    for n, lam, fd, pc in ((8, 1.0, 0, 2), (8, 1.0, 1, 1), (12, 2.0, 0, 1)):
        N = n ** 3
        xg, g = ref.bratu_gmres(n, lam, fd, pc, O.sin_point(N, 0.1), O.sin_point(N, 1.0), 1e-10,
                                 40 if n == 8 else 10, 400)
        tag = f"gmres_n{n}_fd{fd}_pc{pc}"
        arrays[tag] = xg
        arrays[tag + "_counts"] = np.array([g["iters"], g["matvecs"], g["precs"], g["converged"],
                                            g["breakdown"]], np.float64)
    for n, lam, method, pc, rec in ((8, 1.0, 0, 1, 0), (15, 1.0, 1, 2, 1), (16, 6.0, 1, 1, 0)):
        p = O.krylov_params(recycle=rec)
        xs, r = ref.bratu_solve(n, lam, method, pc, 0, p)
        tag = f"solve_n{n}_m{method}_pc{pc}_r{rec}"
        arrays[tag] = xs
        arrays[tag + "_counts"] = np.array([r["iters"], r["n_F"], r["n_JVP"], r["n_D2"],
                                            r["n_lin_iters"], r["n_prec"]], np.float64)
    # tensor-free Chebyshev 2018: curvature on the input of
    # test_tf_chebyshev_2018.cpp:20-25 and solves (method 2)
    for n, lam in ((4, 1.0), (8, 3.0)):
        i = np.arange(n ** 3, dtype=np.float64)
        x = 0.1 + 0.03 * np.sin(0.2 * i)
        d = 0.5 * np.cos(0.17 * i)
        for fd in (0, 1):
            c, h = ref.tf2018_curvature(n, lam, x, d, fd)
            arrays[f"tf2018_curv_n{n}_fd{fd}"] = c
            arrays[f"tf2018_curv_n{n}_fd{fd}_h"] = np.array([h])
    for n, lam, pc, fd in TF_SOLVES:
        p = O.krylov_params(rtol_nl=1e-8, max_nl=20, gmres_restart=30, gmres_max_iters=300)
        xs, r = ref.bratu_solve(n, lam, 2, pc, fd, p)
        tag = f"tf2018_n{n}_l{lam}_pc{pc}_fd{fd}"
        arrays[tag] = xs
        arrays[tag + "_counts"] = np.array([r["converged"], r["iters"], r["n_F"], r["n_JVP"],
                                            r["n_D2"], r["n_lin_iters"], r["n_lin_failures"],
                                            r["n_prec"]], np.float64)
    # inverse_smallest_eigenpair of J(x) with PoissonMG(n) (calibrate_b6.cpp:240-248)
    for n, lam in ((15, 3.0), (7, 0.0)):
        e, v = ref.bratu_eigen_smallest(n, lam, O.sin_point(n ** 3, 0.3))
        tag = f"eig_n{n}_l{lam}"
        arrays[tag] = v
        arrays[tag + "_res"] = np.array([e["converged"], e["eigenvalue"], e["residual_rel"],
                                         e["outer_iterations"], e["linear_iterations"],
                                         e["linear_matvecs"], e["preconditioner_applications"],
                                         e["linear_failures"]], np.float64)
    for N in (8, 24):
        nodes, weights = ref.gauss_legendre_unit(N)
        arrays[f"gl_nodes_{N}"] = nodes
        arrays[f"gl_weights_{N}"] = weights
        x = 1.0 + 0.2 * np.sin(0.37 * np.arange(N) + 0.5)
        d = 0.5 * np.cos(0.23 * np.arange(N) + 0.1)
        arrays[f"chandra_eval_{N}"] = ref.chandra("eval", N, 0.8, x)
        arrays[f"chandra_jvp_{N}"] = ref.chandra("jvp", N, 0.8, x, d)
        arrays[f"chandra_d2_{N}"] = ref.chandra("d2", N, 0.8, x, d)
    np.savez_compressed(os.path.join(HERE, "reference_vectors.npz"), **arrays)
    print(f"wrote {len(data['calibration_n31'])} calibration rows, {len(arrays)} arrays")


if __name__ == "__main__":
    main()
