#!/usr/bin/env python3
"""Bratu3D 256^3 time-to-1e-8 residual on one MI355X.

A step is one complete nonlinear solve of Bratu3D (n = 256 interior nodes per
axis, lambda = 6.5, u0 = 0, ||F|| <= 1e-8 ||F(u0)||) with the device-resident
Newton-Krylov solver: FD J v, GMRES(30) / FGMRES(30), 5-level GMG V(2,2) with
the degree-3 Chebyshev smoother.  Three legs run on the same GPU:
  * fp64:  plain FP64 GMG-GMRES with the reference's forcing, one kernel per
           smoothing step (the speed-up denominator);
  * mixed: FGMRES with FP32 Krylov basis / J v / V-cycle and FP64 F(x),
           difference quotient, Newton update and Givens solve, smoothing
           sweeps temporally blocked (autotuned per level), with the
           final-solve safeguard so that its solution is within 1e-10 of the
           FP64 reference's whatever Krylov path it takes (the value);
  * fp64_fused: the FP64 method with the mixed leg's smoothing kernels and
           the same safeguard (reported next to the value).
`same_policy` times the mixed leg with the reference's forcing and the fp64
leg with the safeguard, so both ratios are on record.  The timed region holds
only solves, inputs resident in HBM.  Rank 0 prints one JSON line.  The CPU
baseline is the FP64 CPU oracle (OpenMP) solving the same problem on the
host; its solutions (reference forcing, and safeguard) are the parity
references of the GPU solutions with the same policy.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "chebyshev-nonlinear-lab_amd")
ORACLE_DIR = os.path.join(ROOT, "oracle")
for _p in (PKG, ORACLE_DIR):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = ("Bratu3D 256³ time-to-1e-8 residual: learned/mixed FGMRES vs plain FP64 GMG")
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def make_params(H, forcing: str, safeguard: bool = False):
    """The reference's KrylovParams with GMRES(30); safeguard: the final-solve
    forcing cap (eta <= 1e-3 once ||F|| <= 1e-6 ||F0||, converged only once the
    last step started there), which makes the final iterate independent of
    the Krylov path: every path lands within ~6e-12 (relative L2) of the
    converged solution at 256^3, so two solvers on different paths agree to
    1e-10, at 4 more FGMRES iterations than the reference's forcing (whose
    final iterate is only ~3e-10 from the converged solution)."""
    extra = dict(tight_res=SAFE_RES, tight_eta=SAFE_ETA) if safeguard else {}
    if forcing == "constant":
        return H.krylov_params(gmres_restart=30, gmres_max_iters=600, eta_init=1e-3,
                               forcing=H.FORCING_CONSTANT, **extra)
    return H.krylov_params(gmres_restart=30, gmres_max_iters=600, **extra)


SAFE_RES, SAFE_ETA = 1e-6, 1e-3


# Coarsest level (16^3 at 256^3 with 5 levels) solved exactly by sine
# transforms in every leg.
COARSE = -1
# Chebyshev smoothing interval [lmax / CHEB_RATIO, lmax] of D^-1 A in every
# leg of the 256^3 workload, the best for the plain FP64 leg and for the
# value.  256^3 lambda = 6.5 on the MI355X (Newton / GMRES iterations):
#   ratio  plain FP64   mixed + safeguard
#     4      5 / 18      5 / 22   15.0 ms
#    12      4 / 15      5 / 21   14.1 ms
#    20      4 / 15      5 / 20   13.4 ms
#    30      4 / 15      5 / 21   14.2 ms
# (the CPU oracle, FP64: 40 -> 4 / 16 plain, 60 -> 5 / 23 with the safeguard).
CHEB_RATIO = 20.0


def leg_configs(H, levels: int, smooth_rows=None, ratio: float = None):
    """fp64: the plain method (one kernel per smoothing step); mixed: FGMRES in
    mixed precision with autotuned temporally blocked smoothing; fp64_fused: the
    FP64 method with the same smoothing kernels as mixed (reported, so that the
    share of the speed-up that comes from precision is visible)."""
    ratio = CHEB_RATIO if ratio is None else ratio
    plain = H.mg_cheb(levels=levels, degree=3, pre=2, post=2, fuse=H.FUSE_NONE,
                      coarse_iters=COARSE, cheb_ratio=ratio)
    if smooth_rows is None:
        fused = H.mg_cheb(levels=levels, degree=3, pre=2, post=2, fuse=H.FUSE_AUTO,
                          coarse_iters=COARSE, cheb_ratio=ratio)
    else:
        fused = H.mg_cheb(levels=levels, degree=3, pre=2, post=2, fuse=H.FUSE_SWEEP,
                          rows=smooth_rows, coarse_iters=COARSE, cheb_ratio=ratio)
    fp64 = H.solver_config(method=H.METHOD_NEWTON, precond=H.PRECOND_GMG,
                           precision=H.PRECISION_FP64, flexible=0, jv_mode=H.DERIV_FD, mg=plain)
    mixed = H.solver_config(method=H.METHOD_NEWTON, precond=H.PRECOND_GMG,
                            precision=H.PRECISION_MIXED, flexible=1, jv_mode=H.DERIV_FD, mg=fused)
    fp64_fused = H.solver_config(method=H.METHOD_NEWTON, precond=H.PRECOND_GMG,
                                 precision=H.PRECISION_FP64, flexible=0, jv_mode=H.DERIV_FD,
                                 mg=fused)
    return {"fp64": fp64, "mixed": mixed, "fp64_fused": fp64_fused}


def time_leg(H, ctx, n, lam, cfg, params, steps, warmup, graphs=False):
    import numpy as np
    s = H.BratuSolver(ctx, n, lam, cfg, restart=30)
    if graphs:
        s.set_graphs(True)
    x = ctx.empty(n ** 3)
    zeros = np.zeros(n ** 3)
    res = None
    for _ in range(warmup):
        x.upload(zeros)
        res = s.solve(params, x)
    times = []
    for _ in range(steps):
        x.upload(zeros)
        ctx.sync()
        t0 = time.perf_counter()
        res = s.solve(params, x)
        ctx.sync()
        times.append(time.perf_counter() - t0)
        log(f"  solve {len(times)}/{steps}: {times[-1]*1e3:.1f} ms, newton={res.iters} "
            f"gmres={res.n_lin_iters} rel={res.res/res.res0:.2e}")
    return s, x, res, times


PROFILE_MD = os.path.join(ROOT, "profiles", "bratu256_lambda6.5_newton_fgmres_gmg_kernels.md")


def fused_grid_threads(n: int, steps: int, rows: int, ichunk: int) -> int:
    """Threads of a k_poly_fused launch (64 x 16*rows regions less 2*steps
    halo per axis, plane chunks, blocks rounded up to a multiple of 8)."""
    out_w, out_h = 64 - 2 * steps, 16 * rows - 2 * steps
    # a region as wide (tall) as the domain covers that axis alone, no halo
    tk = 1 if n <= 64 else -(-n // out_w)
    tj = 1 if n <= 16 * rows else -(-n // out_h)
    tiles = tk * tj * -(-n // ichunk)
    return -(-tiles // 8) * 8 * 1024


def measured_traffic(kernel: str, n: int, ichunk: int, grid: int):
    """HBM bytes per launch of the dominant kernel from the committed
    FETCH_SIZE / WRITE_SIZE profile (calibrated with a copy kernel), if that
    profile is of the same kernel, grid size and launch shape; else None."""
    try:
        with open(PROFILE_MD) as f:
            for line in f:
                if line.startswith("traffic-json:"):
                    t = json.loads(line.split(":", 1)[1])
                    if (t["kernel"] == kernel and t["n"] == n and t["ichunk"] == ichunk
                            and t["grid_threads"] == grid):
                        return int(t["traffic_bytes"])
    except (OSError, ValueError, KeyError):
        pass
    return None


def primitive_table(H, s, reps=20):
    import ctypes as C
    names = {"F": H.PRIM_F, "JVP": H.PRIM_JVP, "PREC": H.PRIM_PREC, "SMOOTH": H.PRIM_SMOOTH,
             "MGS": H.PRIM_MGS, "COPY": H.PRIM_COPY}
    out = {}
    for name, prim in names.items():
        ms, b = C.c_double(), C.c_double()
        H.check(s.ctx.lib.numa_hip_primitive_cost(s.h, prim, reps, C.byref(ms), C.byref(b)))
        out[name] = {"ms": ms.value, "bytes": b.value,
                     "GBs": b.value / (ms.value * 1e-3) / 1e9 if ms.value > 0 else 0.0}
    return out


def cpu_leg(args) -> None:
    """Runs in a child process with OMP_NUM_THREADS set: FP64 CPU oracle solve."""
    import numpy as np
    import numa_hip as H
    import oracle as O
    cfg = leg_configs(H, args.levels, ratio=args.cheb_ratio)["fp64"]
    m = O.MGConfig(**{k: getattr(cfg.mg, k) for k, _ in cfg.mg._fields_})
    method = H.METHOD_CHEBYSHEV if args.cpu_method == "cheb" else cfg.method
    ocfg = O.SolverConfig(method=method, precond=cfg.precond, precision=0, flexible=0,
                          jv_mode=cfg.jv_mode, d2_mode=cfg.d2_mode, mg=m)
    p = make_params(H, args.forcing, safeguard=args.cpu_safeguard)
    op = O.KrylovParams(**{k: getattr(p, k) for k, _ in p._fields_})
    o = O.Oracle()
    t0 = time.perf_counter()
    x, r = o.bratu_solve(args.n, args.lam, ocfg, op)
    t = time.perf_counter() - t0
    np.save(args.cpu_out, x)
    print(json.dumps({"seconds": t, "threads": o.threads(), "result": O.result_dict(r)}))


def run_cpu_baseline(args, safeguard: bool = False, method: str = "newton"):
    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
               OMP_PLACES="cores")
    import tempfile
    out = os.path.join(tempfile.gettempdir(),
                       f"numa_cpu_solution_{os.getpid()}_{int(safeguard)}_{method}.npy")
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-leg", "--n", str(args.n), "--lam",
           str(args.lam), "--levels", str(args.levels), "--forcing", args.forcing,
           "--cpu-out", out, "--cheb-ratio", str(args.cheb_ratio), "--cpu-method", method] + (
               ["--cpu-safeguard"] if safeguard else [])
    log(f"cpu baseline: FP64 oracle {method} solve at n={args.n} with {threads} threads"
        f"{' (final-solve safeguard)' if safeguard else ''}")
    p = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=args.cpu_timeout)
    if p.returncode != 0:
        raise RuntimeError("cpu leg failed: " + p.stderr[-2000:])
    info = json.loads(p.stdout.strip().splitlines()[-1])
    return info, out


def chebyshev_outer_leg(H, ctx, args, params, psafe) -> dict:
    """The reference's Chebyshev outer iteration (chebyshev_inexact,
    krylov_newton.hpp:116-189: per step J z = F and J u = F''[z,z] with the
    central-difference F'', x+ = x - z - u/2 when it lowers ||F||) on the same
    256^3 workload: the plain FP64 method (reference forcing) and the mixed
    one (safeguard), timed as the Newton legs, each against the CPU oracle's
    Chebyshev solve with the same policy."""
    import numpy as np
    cfgs = leg_configs(H, args.levels, args.smooth_rows, args.cheb_ratio)
    out = {"workload": f"bratu3d_{args.n}^3_lambda{args.lam}_chebyshev_inexact_fd_d2_"
                       f"gmres30_gmg{args.levels}_cheb3_v22"}
    sols = {}
    for name, pp in (("fp64", params), ("mixed", psafe)):
        c0 = cfgs[name]
        cfg = H.solver_config(method=H.METHOD_CHEBYSHEV, precond=c0.precond,
                              precision=c0.precision, flexible=c0.flexible, jv_mode=c0.jv_mode,
                              d2_mode=H.DERIV_FD, mg=c0.mg)
        s, x, r, t = time_leg(H, ctx, args.n, args.lam, cfg, pp, args.steps, args.warmup)
        out[name] = {"s": round(float(np.mean(t)), 6), "outer_iters": int(r.iters),
                     "gmres_iters": int(r.n_lin_iters), "n_F": int(r.n_F),
                     "n_JVP": int(r.n_JVP), "n_D2": int(r.n_D2),
                     "converged": int(r.converged), "final_rel_residual": r.res / r.res0,
                     "policy": "reference forcing" if pp is params else "safeguard"}
        sols[name] = x.numpy()
        s.close()
        x.free()
        log(f"chebyshev outer {name}: {out[name]}")
    out["speedup_mixed_vs_fp64"] = round(out["fp64"]["s"] / out["mixed"]["s"], 3)
    if not args.no_cpu:
        for name, safe in (("fp64", False), ("mixed", True)):
            info, path = run_cpu_baseline(args, safeguard=safe, method="cheb")
            xc = np.load(path)
            os.remove(path)
            res = info["result"]
            out[name]["cpu_fp64"] = {"s": round(info["seconds"], 3), "cores": info["threads"],
                                     "outer_iters": res["iters"], "n_D2": res["n_D2"],
                                     "gmres_iters": res["n_lin_iters"]}
            out[name]["rel_l2_vs_cpu_fp64"] = float(np.linalg.norm(sols[name] - xc) /
                                                    np.linalg.norm(xc))
    return out


CFG0_N, CFG0_LAM = 32, 6.0


def config0_cfg(H):
    """32^3, lambda = 6: FP64 Newton, FD J v, GMRES(30) + V(2,2) weighted-Jacobi
    multigrid (full even-n hierarchy, exact coarsest solve), constant forcing 1e-3."""
    return H.solver_config(method=H.METHOD_NEWTON, precond=H.PRECOND_GMG,
                           precision=H.PRECISION_FP64, flexible=0, jv_mode=H.DERIV_FD,
                           mg=H.mg_wjacobi(coarse_iters=-1))


def config0_cpu_leg(args) -> None:
    """Child process: the oracle's FP64 solve of the 32^3 plain-method
    workload, one warm-up then `reps` timed solves."""
    import numpy as np
    import numa_hip as H
    import oracle as O
    cfg = config0_cfg(H)
    m = O.MGConfig(**{k: getattr(cfg.mg, k) for k, _ in cfg.mg._fields_})
    ocfg = O.SolverConfig(method=cfg.method, precond=cfg.precond, precision=0, flexible=0,
                          jv_mode=cfg.jv_mode, d2_mode=cfg.d2_mode, mg=m)
    p = make_params(H, "constant")
    op = O.KrylovParams(**{k: getattr(p, k) for k, _ in p._fields_})
    o = O.Oracle()
    o.bratu_solve(CFG0_N, CFG0_LAM, ocfg, op)  # warm-up
    times = []
    for _ in range(args.cfg0_reps):
        t0 = time.perf_counter()
        x, r = o.bratu_solve(CFG0_N, CFG0_LAM, ocfg, op)
        times.append(time.perf_counter() - t0)
    np.save(args.cpu_out, x)
    print(json.dumps({"times": times, "threads": o.threads(), "result": O.result_dict(r)}))


def config0_leg(H, ctx, args) -> dict:
    """The reference's plain method at 32^3 on the GPU (median of 5 solves after
    a warm-up) and the same FP64 solve by the CPU oracle (median of 5 after a
    warm-up, OpenMP), with the counters and the solution compared."""
    import numpy as np
    cfg = config0_cfg(H)
    p = make_params(H, "constant")
    s = H.BratuSolver(ctx, CFG0_N, CFG0_LAM, cfg, restart=30)
    x = ctx.empty(CFG0_N ** 3)
    zeros = np.zeros(CFG0_N ** 3)
    x.upload(zeros)
    s.solve(p, x)
    times = []
    for _ in range(5):
        x.upload(zeros)
        ctx.sync()
        t0 = time.perf_counter()
        r = s.solve(p, x)
        ctx.sync()
        times.append(time.perf_counter() - t0)
    xg = x.numpy()
    s.close()
    x.free()
    keys = ("iters", "n_F", "n_JVP", "n_lin_iters", "n_prec")
    out = {"workload": "bratu3d_32^3_lambda6.0_newton_gmres30_gmg_wjacobi_v22_fd_jv_forcing1e-3",
           "gpu": {"median_s": round(float(np.median(times)), 6), "solves": 5,
                   **{k: int(getattr(r, k)) for k in keys}}}
    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
               OMP_PLACES="cores")
    import tempfile
    path = os.path.join(tempfile.gettempdir(), f"numa_cfg0_{os.getpid()}.npy")
    pr = subprocess.run([sys.executable, os.path.abspath(__file__), "--config0-cpu-leg",
                         "--cpu-out", path, "--cfg0-reps", "5"], env=env, capture_output=True,
                        text=True, timeout=args.cpu_timeout)
    if pr.returncode == 0:
        info = json.loads(pr.stdout.strip().splitlines()[-1])
        xc = np.load(path)
        os.remove(path)
        out["cpu_baseline"] = {"median_s": round(float(np.median(info["times"])), 6),
                               "solves": len(info["times"]), "warmup": 1,
                               "cores": info["threads"], "kind": "port",
                               "cpu_model": cpu_model(),
                               **{k: int(info["result"][k]) for k in keys}}
        out["counters_identical"] = all(out["gpu"][k] == out["cpu_baseline"][k] for k in keys)
        out["rel_l2_gpu_vs_cpu"] = float(np.linalg.norm(xg - xc) / np.linalg.norm(xc))
        out["speedup_gpu_vs_cpu"] = round(out["cpu_baseline"]["median_s"] /
                                          out["gpu"]["median_s"], 2)
    log(f"config 0: {out}")
    return out


CH_N, CH_B, CH_SEED = 8192, 256, 20261203
FP64_MFMA_PEAK_TFLOPS = 78.6
BF16_MFMA_PEAK_TFLOPS = 2500.0


def chandra_cs():
    """256 values of c uniform in [0.5, 0.9999] from a recorded seed
    (mt19937-independent: numpy PCG64)."""
    import numpy as np
    return np.random.default_rng(CH_SEED).uniform(0.5, 0.9999, CH_B)


def chandra_leg(H, ctx, args) -> dict:
    """Chandrasekhar H-equation, midpoint rule, Kelley form, N = 8192, a batch of
    256 c values, x0 = 1, Newton-GMRES(20), tol 1e-10: FP64 vs split-BF16 J v."""
    import numpy as np
    cs = chandra_cs()
    p = H.krylov_params(rtol_nl=1e-10, gmres_restart=20, gmres_max_iters=400)
    out = {"workload": f"chandrasekhar_midpoint_kelley_N{CH_N}_batch{CH_B}_newton_gmres20_tol1e-10",
           "c_seed": CH_SEED}
    # the device's measured dense MFMA throughput, beside the published peaks
    peak64, peak16 = ctx.mfma_peak(1), ctx.mfma_peak(0)
    out["measured_mfma_tflops"] = {"fp64_16x16x4": round(peak64, 1),
                                   "bf16_32x32x16": round(peak16, 1)}
    sols = {}
    for name, prec in (("fp64", H.PRECISION_FP64), ("mixed", H.PRECISION_MIXED)):
        ch = H.ChandrasekharBatch(ctx, CH_N, cs, rule=1, form=1, precision=prec, restart=20)
        x = ctx.array(np.ones(CH_B * CH_N))
        ch.solve_device(p, x)  # warm-up
        x.upload(np.ones(CH_B * CH_N))
        ctx.sync()
        t0 = time.perf_counter()
        res = ch.solve_device(p, x)
        ctx.sync()
        t = time.perf_counter() - t0
        sols[name] = x.numpy().reshape(CH_B, CH_N)
        ms64, fl = ch.gemm_cost(H.PRECISION_FP64, 10)
        leg = {"solve_s": round(t, 5), "converged": int(sum(r.converged for r in res)),
               "max_newton": int(max(r.iters for r in res)),
               "total_jv": int(sum(r.n_JVP for r in res)),
               "gemm_fp64_ms": round(ms64, 4),
               "gemm_fp64_tflops": round(fl / (ms64 * 1e-3) / 1e12, 2),
               "gemm_fp64_frac_of_78.6": round(fl / (ms64 * 1e-3) / 1e12 / FP64_MFMA_PEAK_TFLOPS, 4),
               "gemm_fp64_frac_of_measured": round(fl / (ms64 * 1e-3) / 1e12 / peak64, 4)}
        if prec == H.PRECISION_MIXED:
            msb, _ = ch.gemm_cost(H.PRECISION_MIXED, 10)
            leg.update({"gemm_bf16x3_ms": round(msb, 4),
                        "gemm_bf16x3_useful_tflops": round(fl / (msb * 1e-3) / 1e12, 2),
                        "gemm_bf16x3_issued_tflops": round(3 * fl / (msb * 1e-3) / 1e12, 2),
                        "gemm_bf16x3_issued_frac_of_2500": round(
                            3 * fl / (msb * 1e-3) / 1e12 / BF16_MFMA_PEAK_TFLOPS, 4),
                        "gemm_bf16x3_issued_frac_of_measured": round(
                            3 * fl / (msb * 1e-3) / 1e12 / peak16, 4)})
        out[name] = leg
        log(f"chandrasekhar {name}: {t*1e3:.1f} ms, converged {leg['converged']}/{CH_B}, "
            f"max newton {leg['max_newton']}, {leg}")
        ch.close()
    out["speedup_mixed_vs_fp64"] = round(out["fp64"]["solve_s"] / out["mixed"]["solve_s"], 3)
    out["mixed_vs_fp64_max_rel_l2"] = float(max(
        np.linalg.norm(sols["mixed"][b] - sols["fp64"][b]) / np.linalg.norm(sols["fp64"][b])
        for b in range(CH_B)))
    return out


LEARN_LAMBDAS = (5.0, 6.0, 6.5, 6.7)
DEFAULT_WEIGHTS = os.path.join(PKG, "weights", "convnet_200k_n16-32")
CONV_FLOP_PER_POINT = 2 * 27 * (1 * 24 + 4 * 24 * 24 + 24 * 1)  # 127,008


def learned_leg(H, ctx, args, params) -> dict:
    """Learned preconditioner.  The conv net is trained on the GPU on
    (residual, V(2,2) correction) pairs generated by mixed Newton-FGMRES solves
    (lambda ~ U[1, 6.7], residual-norm loss).  By default the parameters are
    the ones train_convnet.py wrote for 200k pairs on 16^3-32^3 grids
    (--learned-weights; the 64^3-128^3 grids at the same budget train a worse
    network); --learned-retrain trains here with --learned-pairs pairs on
    --learned-sizes grids instead.  Then Newton-FGMRES(30) at n^3 for lambda in {5.0, 6.0,
    6.5, 6.7} runs with the network and the V-cycle fallback (a step reducing
    the residual by less than 0.5 switches the rest of that linear solve to the
    V-cycle) next to the V-cycle alone, both with the final-solve safeguard so
    that their solutions can be compared at 1e-10."""
    import numpy as np
    n = args.n
    net = H.ConvNet(ctx, seed=21)
    wpath = args.learned_weights
    if wpath and os.path.exists(wpath + ".npy") and not args.learned_retrain:
        # parameters trained on the GPU by train_convnet.py at the full budget
        net.set_params(np.load(wpath + ".npy"))
        rec = json.load(open(wpath + ".json")) if os.path.exists(wpath + ".json") else {}
        out = {"train": {k: rec.get(k) for k in ("pairs", "steps", "solves", "seconds", "sizes",
                                                 "batch", "lr", "loss", "loss_first_window",
                                                 "loss_last_window")}}
        out["train"]["source"] = os.path.relpath(wpath, ROOT) + ".npy (train_convnet.py)"
        log(f"learned: parameters from {wpath}.npy ({rec.get('pairs')} pairs)")
    else:
        sizes = tuple(int(v) for v in args.learned_sizes.split(","))
        tcfg = H.net_train_config(pairs=args.learned_pairs, batch=4, sizes=sizes, lr=2e-3,
                                  seed=5, mg=H.mg_cheb(levels=3), loss=H.LOSS_RESIDUAL)
        st, hist = net.train_solver(tcfg, hist_len=1000000)
        k = max(1, len(hist) // 10)
        out = {"train": {"pairs": int(st.pairs), "steps": int(st.steps), "solves": int(st.solves),
                         "seconds": round(st.seconds, 3), "sizes": list(sizes),
                         "loss": "residual", "loss_first_window": round(float(hist[:k].mean()), 4),
                         "loss_last_window": round(float(hist[-k:].mean()), 4),
                         "source": "trained in this run"}}
        log(f"learned: trained on {st.pairs} pairs in {st.seconds:.2f} s, loss "
            f"{out['train']['loss_first_window']} -> {out['train']['loss_last_window']}")
    if args.learned_save:
        np.save(args.learned_save, net.get_params())
    params = make_params(H, args.forcing, safeguard=True)
    # the value leg's V-cycle (smoothing interval, autotuned fused smoothing)
    mg = leg_configs(H, args.levels, args.smooth_rows, args.cheb_ratio)["mixed"].mg
    cfg_l = H.solver_config(method=H.METHOD_NEWTON, precond=H.PRECOND_LEARNED,
                            precision=H.PRECISION_MIXED, flexible=1, jv_mode=H.DERIV_FD, mg=mg)
    cfg_g = H.solver_config(method=H.METHOD_NEWTON, precond=H.PRECOND_GMG,
                            precision=H.PRECISION_MIXED, flexible=1, jv_mode=H.DERIV_FD, mg=mg)
    zeros = np.zeros(n ** 3)
    runs = []
    for lam in LEARN_LAMBDAS:
        row = {"lambda": lam}
        sols = {}
        for name, cfg in (("learned", cfg_l), ("gmg", cfg_g)):
            s = H.BratuSolver(ctx, n, lam, cfg, restart=30)
            if name == "learned":
                s.set_net(net, 0.5)
            x = ctx.empty(n ** 3)
            x.upload(zeros)
            s.solve(params, x)  # warm-up
            x.upload(zeros)
            st0 = s.learned_stats() if name == "learned" else None
            ctx.sync()
            t0 = time.perf_counter()
            r = s.solve(params, x)
            ctx.sync()
            t = time.perf_counter() - t0
            row[name] = {"s": round(t, 5), "newton": int(r.iters), "fgmres": int(r.n_lin_iters),
                         "converged": int(r.converged), "rel_res": r.res / r.res0}
            if name == "learned":
                st1 = s.learned_stats()
                row[name].update({k2: st1[k2] - st0[k2] for k2 in st1})
            sols[name] = x.numpy()
            s.close()
            x.free()
        row["rel_l2_learned_vs_gmg"] = float(np.linalg.norm(sols["learned"] - sols["gmg"]) /
                                             np.linalg.norm(sols["gmg"]))
        ref = getattr(args, "_safe_solution", None)
        if ref is not None and lam == args.lam and ref.size == n ** 3:
            row["rel_l2_learned_vs_cpu_fp64"] = float(np.linalg.norm(sols["learned"] - ref) /
                                                      np.linalg.norm(ref))
        log(f"  lambda {lam}: {row}")
        runs.append(row)
    out["solves"] = runs
    # what the trained network gains, network only (no fallback): FGMRES
    # residual after 8 steps on J(0) z = b at 64^3
    gain = {}
    gn = 64
    for rhs in ("random", "constant"):
        bv = (np.random.default_rng(31).normal(size=gn ** 3) if rhs == "random"
              else -6.0 * np.ones(gn ** 3))
        b, x0, z = ctx.array(bv), ctx.array(np.zeros(gn ** 3)), ctx.empty(gn ** 3)
        row = {}
        untrained = H.ConvNet(ctx, seed=21)
        for name, pc, nt in (("trained", H.PRECOND_LEARNED, net),
                             ("untrained", H.PRECOND_LEARNED, untrained),
                             ("none", H.PRECOND_NONE, None), ("gmg", H.PRECOND_GMG, None)):
            c = H.solver_config(precond=pc, precision=H.PRECISION_MIXED, flexible=1,
                                mg=H.mg_cheb(levels=4))
            s = H.BratuSolver(ctx, gn, 6.0, c, restart=30)
            if nt is not None:
                s.set_net(nt, 1.0)
            row[name] = s.gmres(x0, b, z, 1e-14, 30, 8).rel_res
            s.close()
        untrained.close()
        gain[rhs] = row
    out["gain_64cubed_rel_res_after_8_fgmres_steps"] = gain
    log(f"  gain: {gain}")
    ms_inf = net.cost(n, 0, 1, 3)
    ms_tr = net.cost(64, 1, 4, 3)
    out["inference_ms"] = round(ms_inf, 3)
    out["inference_useful_tflops"] = round(CONV_FLOP_PER_POINT * n ** 3 / (ms_inf * 1e-3) / 1e12, 1)
    out["inference_frac_of_bf16_2500"] = round(out["inference_useful_tflops"] /
                                               BF16_MFMA_PEAK_TFLOPS, 4)
    peak16 = ctx.mfma_peak(0)
    out["measured_bf16_mfma_tflops"] = round(peak16, 1)
    out["inference_frac_of_measured_bf16"] = round(out["inference_useful_tflops"] / peak16, 4)
    out["train_step_64cubed_batch4_ms"] = round(ms_tr, 3)
    net.close()
    return out


SWEEP_SEED, SWEEP_COUNT, SWEEP_SIZES = 20261203, 96, (32, 48, 64, 96)


def sweep_leg(H, ctx, args, params, batch_only: bool = False) -> dict:
    """Batched sweep: 96 Bratu3D instances, n from {32,48,64,96} and lambda ~
    U[1.0, 6.7] drawn with std::mt19937_64(20261203), u0 = 0.1 U[0,1) from
    mt19937_64(seed*7919 + 13 + i); mixed Newton-FGMRES(30) + GMG (the main
    leg's configuration: smoothing interval ratio 20, autotuned fused
    smoothing).  One eager worker (the
    instances one after another) next to W workers with captured Arnoldi
    steps; each configuration is run once untimed (allocation, tuning,
    capture) and once timed.  Then the batched solver: all 96 instances in
    one launch per operation over the instance table and one graph launch per
    lockstep Arnoldi step (best of two timed solves after a tuning one), with
    its algorithmic HBM bytes, and a sample checked bit for bit against the
    single-instance solver with unfused smoothing."""
    import numpy as np
    n, lam = H.sweep_instances(SWEEP_SEED, SWEEP_COUNT, sizes=SWEEP_SIZES)
    x0 = [H.uniform_host(SWEEP_SEED * 7919 + 13 + i, int(n[i]) ** 3, 0.1)
          for i in range(SWEEP_COUNT)]
    xs = [ctx.array(x) for x in x0]
    cfg = H.solver_config(method=H.METHOD_NEWTON, precond=H.PRECOND_GMG,
                          precision=H.PRECISION_MIXED, flexible=1, jv_mode=H.DERIV_FD,
                          mg=H.mg_cheb(levels=5, fuse=H.FUSE_AUTO, coarse_iters=COARSE,
                                       cheb_ratio=CHEB_RATIO))
    out = {"workload": f"bratu3d_sweep_{SWEEP_COUNT}_instances_n{'-'.join(map(str, SWEEP_SIZES))}"
                       f"_lambda1.0-6.7_u0_0.1noise_mixed_newton_fgmres30_gmg",
           "seed": SWEEP_SEED, "sizes": {str(k): int((n == k).sum()) for k in SWEEP_SIZES}}
    sols = None
    for workers, graphs in (() if batch_only else ((1, False), (4, True), (8, True))):
        sw = H.Sweep(ctx, cfg, workers=workers, restart=30)
        for rep_ in range(2):
            for x, v in zip(xs, x0):
                x.upload(v)
            res, sec = sw.run(n, lam, xs, params, graphs=graphs)
        sw.close()
        key = f"workers{workers}_{'graphs' if graphs else 'eager'}"
        out[key] = {"seconds": round(sec, 4), "solves_per_s": round(SWEEP_COUNT / sec, 1),
                    "converged": int(sum(r.converged for r in res)),
                    "newton_total": int(sum(r.iters for r in res)),
                    "fgmres_total": int(sum(r.n_lin_iters for r in res))}
        cur = [x.numpy() for x in xs]
        if sols is None:
            sols = cur
        else:
            out[key]["max_abs_diff_vs_workers1"] = float(max(np.max(np.abs(a - b))
                                                             for a, b in zip(cur, sols)))
        log(f"sweep {key}: {out[key]}")
    # the batched solver: one launch per operation over the instance table and
    # one graph launch per lockstep Arnoldi step for all instances
    b = H.Batch(ctx, n, lam, cfg, restart=30)
    b.set_graphs(True)
    times = []
    for rep_ in range(3):  # tuning and capture, then two timed solves
        for x, v in zip(xs, x0):
            x.upload(v)
        ctx.sync()
        b.stats(reset=True)
        res, sec = b.solve(xs, params)
        st = b.stats()
        times.append((sec, st))
    sec, st = min(times[1:], key=lambda t: t[0])
    b.close()
    cur = [x.numpy() for x in xs]
    out["batched_graphs"] = {
        "seconds": round(sec, 4), "solves_per_s": round(SWEEP_COUNT / sec, 1),
        "converged": int(sum(r.converged for r in res)),
        "newton_total": int(sum(r.iters for r in res)),
        "fgmres_total": int(sum(r.n_lin_iters for r in res)),
        "lockstep_arnoldi_steps": st["arnoldi_steps"], "graph_launches": st["graph_launches"],
        "other_launches": st["launches"], "host_syncs": st["host_syncs"],
        "algorithmic_bytes": st["bytes"], "achieved_GBs": round(st["bytes"] / sec / 1e9, 1),
        "frac_of_8000": round(st["bytes"] / sec / 1e9 / HBM_PEAK_GBS, 4),
        "max_abs_diff_vs_workers1": None if sols is None else
        float(max(np.max(np.abs(a - b_)) for a, b_ in zip(cur, sols)))}
    # bit for bit the single-instance solver with unfused smoothing (sample)
    alone = H.solver_config(method=H.METHOD_NEWTON, precond=H.PRECOND_GMG,
                            precision=H.PRECISION_MIXED, flexible=1, jv_mode=H.DERIV_FD,
                            mg=H.mg_cheb(levels=5, fuse=H.FUSE_NONE, coarse_iters=COARSE,
                                         cheb_ratio=CHEB_RATIO))
    same = []
    for k in SWEEP_SIZES:
        i = int(np.flatnonzero(n == k)[0])
        s = H.BratuSolver(ctx, int(n[i]), float(lam[i]), alone, restart=30)
        xi = x0[i].copy()
        s.solve_host(params, xi)
        s.close()
        same.append(bool(np.array_equal(xi, cur[i])))
    out["batched_graphs"]["bitwise_vs_alone_unfused_sample"] = same
    log(f"sweep batched_graphs: {out['batched_graphs']}")
    best = max((v["solves_per_s"], k) for k, v in out.items() if isinstance(v, dict) and
               "solves_per_s" in v)
    out["best"] = {"config": best[1], "solves_per_s": best[0]}
    for x in xs:
        x.free()
    return out


def chandra_cpu_sample(args) -> None:
    """Child process: FP64 oracle Newton-GMRES(20) for the first 2 c values."""
    import oracle as O
    cs = chandra_cs()[:2]
    o = O.Oracle()
    cfg = O.solver_config(method=0, precond=O.PRECOND_NONE, jv_mode=O.DERIV_ANALYTIC)
    p = O.krylov_params(rtol_nl=1e-10, gmres_restart=20, gmres_max_iters=400)
    t0 = time.perf_counter()
    for c in cs:
        o.chandra_solve(CH_N, float(c), cfg, p, rule=1, form=1)
    t = time.perf_counter() - t0
    print(json.dumps({"seconds": t, "instances": len(cs), "threads": o.threads()}))


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--lam", type=float, default=6.5)
    ap.add_argument("--levels", type=int, default=5)
    ap.add_argument("--forcing", choices=["ew", "constant"], default="ew")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-timeout", type=float, default=240.0)
    ap.add_argument("--legs", default="fp64,mixed,fp64_fused")
    ap.add_argument("--graphs", action="store_true",
                    help="replay the Arnoldi steps of the mixed and fp64_fused legs from hipGraphs")
    ap.add_argument("--no-chandra", action="store_true", help="skip the Chandrasekhar leg")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-preconditioner leg")
    ap.add_argument("--learned-pairs", type=int, default=4000)
    ap.add_argument("--learned-sizes", default="16,24,32",
                    help="training grid sizes for --learned-retrain")
    ap.add_argument("--learned-save", default="", help="write the trained parameters (.npy)")
    ap.add_argument("--learned-weights", default=DEFAULT_WEIGHTS,
                    help="parameters (path without .npy) to use instead of training")
    ap.add_argument("--learned-retrain", action="store_true",
                    help="train with --learned-pairs / --learned-sizes even if weights exist")
    ap.add_argument("--no-sweep", action="store_true", help="skip the batched-sweep leg")
    ap.add_argument("--smooth-rows", type=int, default=None,
                    help="pin the fused legs' smoothing to one launch per sweep with this "
                         "rows-per-thread value (2 or 4 for FP32 march, -1..-3 box) on every "
                         "level, plane chunks from the wave-filling model, instead of autotuning")
    ap.add_argument("--only", choices=["chandra", "learned", "sweep", "batch"], default=None,
                    help="run only this leg and print its JSON (for profiling); batch: the "
                         "sweep's batched solver alone")
    ap.add_argument("--no-config0", action="store_true", help="skip the 32^3 plain-method leg")
    ap.add_argument("--no-chebyshev", action="store_true",
                    help="skip the Chebyshev outer-iteration legs")
    ap.add_argument("--no-safeguard", action="store_true",
                    help="skip the legs timed with the other forcing policy")
    ap.add_argument("--cheb-ratio", type=float, default=CHEB_RATIO,
                    help="Chebyshev smoothing interval [lmax/ratio, lmax] of the 256^3 legs")
    ap.add_argument("--reference-forcing", action="store_true",
                    help="time every leg with the reference's forcing (no final-solve "
                         "safeguard); parity then holds only along the CPU's Krylov path")
    ap.add_argument("--cpu-leg", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-safeguard", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-method", default="newton", help=argparse.SUPPRESS)
    ap.add_argument("--config0-cpu-leg", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--cfg0-reps", type=int, default=5, help=argparse.SUPPRESS)
    ap.add_argument("--chandra-cpu-leg", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-out", default="", help=argparse.SUPPRESS)
    args = ap.parse_args()

    if args.cpu_leg:
        cpu_leg(args)
        return 0
    if args.chandra_cpu_leg:
        chandra_cpu_sample(args)
        return 0
    if args.config0_cpu_leg:
        config0_cpu_leg(args)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (this build "
              "runs the solver on a single MI355X; multi-GPU sharding is design only)",
              file=sys.stderr)
        return 2

    import numpy as np
    import numa_hip as H

    H.load()
    ctx = H.Context(0)
    params = make_params(H, args.forcing)
    if args.only:
        part = {"chandra": lambda: chandra_leg(H, ctx, args),
                "learned": lambda: learned_leg(H, ctx, args, params),
                "sweep": lambda: sweep_leg(H, ctx, args, params),
                "batch": lambda: sweep_leg(H, ctx, args, params, batch_only=True)}[args.only]()
        print(json.dumps({args.only: part}))
        ctx.close()
        return 0
    cfgs = leg_configs(H, args.levels, args.smooth_rows, args.cheb_ratio)
    psafe = make_params(H, args.forcing, safeguard=not args.reference_forcing)
    leg_params = {"fp64": params, "mixed": psafe, "fp64_fused": psafe}
    legs = {}
    for name in args.legs.split(","):
        log(f"leg {name}: n={args.n} lambda={args.lam}")
        s, x, res, times = time_leg(H, ctx, args.n, args.lam, cfgs[name], leg_params[name],
                                    args.steps, args.warmup,
                                    graphs=args.graphs and name != "fp64")
        legs[name] = {"solver": s, "x": x, "res": res, "times": times,
                      "prims": primitive_table(H, s), "fusion": s.mg_fusion()}
        log(f"  primitives: " + ", ".join(f"{k} {v['ms']:.3f} ms {v['GBs']:.0f} GB/s"
                                          for k, v in legs[name]["prims"].items()))
    main_leg = "mixed" if "mixed" in legs else list(legs)[0]
    L = legs[main_leg]
    t_main = float(np.mean(L["times"]))
    r = L["res"]

    # roofline of the dominant kernel: the finest-level smoothing launch
    sm = L["prims"]["SMOOTH"]
    f0 = L["fusion"][0] if L["fusion"] else (H.FUSE_NONE, 0, 0)
    traffic = None
    kbase = None
    if f0[0] == H.FUSE_NONE:
        kname = "k_stencil_lds<EpPoly> (finest Chebyshev smoothing step)"
    elif f0[2] < 0:  # box form
        steps = 3
        ei, ej, ek = H.BOX_EDGES[int(f0[2])]
        kbase = (f"k_poly_box<{'float' if main_leg == 'mixed' else 'double'},{steps},"
                 f"{ei},{ej},{ek}>")
        kname = f"{kbase} (finest-level fused Chebyshev sweep, box form)"
        boxes = -(-args.n // ei) * -(-args.n // ej) * -(-args.n // ek)
        traffic = measured_traffic(kbase, args.n, 0, -(-boxes // 8) * 8 * 512)
    else:
        steps = 3 if f0[0] == H.FUSE_SWEEP else 6
        kbase = f"k_poly_fused<{'float' if main_leg == 'mixed' else 'double'},{steps},{f0[2]}>"
        kname = (f"{kbase} (finest-level fused Chebyshev "
                 f"{'sweep' if f0[0] == H.FUSE_SWEEP else 'phase'}, ichunk {f0[1]})")
        traffic = measured_traffic(kbase, args.n, int(f0[1]),
                                   fused_grid_threads(args.n, steps, int(f0[2]), int(f0[1])))
    copy = L["prims"]["COPY"]
    # the dominant kernel's average duration inside a solve of the timed
    # workload: HIP events around each finest-level sweep that reads x and b
    # (one more solve of the same leg, eager)
    in_solve_ms, in_solve_n = None, 0
    if f0[0] != H.FUSE_NONE:
        si = H.BratuSolver(ctx, args.n, args.lam, cfgs[main_leg], restart=30)
        si.set_sweep_events(True)
        xi = ctx.empty(args.n ** 3)
        xi.upload(np.zeros(args.n ** 3))
        si.solve(leg_params[main_leg], xi)
        in_solve_ms, in_solve_n = si.sweep_time()
        si.close()
        xi.free()
    avg_ms = in_solve_ms if in_solve_n else sm["ms"]
    ach = sm["bytes"] / (avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": kname,
                "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
                "measured_copy_GBs": round(copy["GBs"], 1),
                "frac_of_measured_copy": round(ach / copy["GBs"], 4) if copy["GBs"] else None,
                "measured_copy_kernel": "k_convert<true,double,double>, 2^24 doubles (read + "
                                        "write 268 MB), HIP events, 20 launches",
                "traffic_source": ("profiles/" + os.path.basename(PROFILE_MD) + " (FETCH_SIZE + "
                                   "WRITE_SIZE per launch, copy-kernel calibrated)")
                if traffic is not None else None,
                "algorithmic_bytes_per_launch": sm["bytes"], "avg_launch_ms": avg_ms,
                "avg_launch_source": (f"start / stop events of the {in_solve_n} finest-level "
                                      "sweeps that read x and b in one solve of the timed leg, "
                                      "stamped by their dispatch (hipExtLaunchKernelGGL)")
                if in_solve_n else "HIP events around 20 back-to-back launches",
                "avg_launch_ms_back_to_back": sm["ms"],
                "frac_back_to_back": round(sm["GBs"] / HBM_PEAK_GBS, 4)}
    if kbase == "k_poly_fused<float,3,2>":
        # counted in the AMD assembly of this kernel's main loop
        roofline["limiter"] = ("VALU issue and the per-plane barrier: 61 VALU instructions "
                               "per plane per thread for 6 node updates (~24 us of issue; packed "
                               "FP32 is not used), halo columns, rows and planes recomputed "
                               "(~1.7x the useful updates), one workgroup barrier per plane "
                               "(~9 us); the plane loads ~6 us (measured on copies of the kernel "
                               "without them)")

    out = {"metric": METRIC, "value": round(t_main, 6), "unit": "s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(t_main * 1e3, 3),
           "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
           "dtype": "f32+f64 (mixed)" if main_leg == "mixed" else "f64",
           "data": f"synthetic: u0 = 0, lambda = {args.lam}",
           "config": {"workload": f"bratu3d_{args.n}^3_lambda{args.lam}_newton_"
                      f"{'fgmres' if main_leg == 'mixed' else 'gmres'}30_gmg{args.levels}"
                      "_cheb3_v22_fd_jv", "n": args.n, "lambda": args.lam,
                      "tol": 1e-8, "forcing": args.forcing, "levels": args.levels,
                      "final_solve_safeguard": None if args.reference_forcing else
                      {"tight_res": SAFE_RES, "tight_eta": SAFE_ETA},
                      "coarse_solve": "exact (sine transforms)" if COARSE < 0 else
                      f"{COARSE} Chebyshev steps"},
           "newton_iters": int(r.iters), "gmres_iters": int(r.n_lin_iters),
           "n_F": int(r.n_F), "n_JVP": int(r.n_JVP), "n_prec": int(r.n_prec),
           "final_rel_residual": r.res / r.res0, "roofline": roofline,
           "primitives": {k: {kk: round(vv, 5) for kk, vv in v.items()}
                          for k, v in L["prims"].items()}}
    out["mg_fusion_per_level"] = [list(f) for f in L["fusion"]]
    if "fp64" in legs and main_leg != "fp64":
        t64 = float(np.mean(legs["fp64"]["times"]))
        r64 = legs["fp64"]["res"]
        out["fp64_gmg_gmres_s"] = round(t64, 6)
        out["speedup_vs_fp64"] = round(t64 / t_main, 3)
        out["fp64_newton_iters"] = int(r64.iters)
        out["fp64_gmres_iters"] = int(r64.n_lin_iters)
    if "fp64_fused" in legs and main_leg != "fp64_fused":
        tf = float(np.mean(legs["fp64_fused"]["times"]))
        out["fp64_fused_smoothing_s"] = round(tf, 6)
        out["speedup_vs_fp64_fused_smoothing"] = round(tf / t_main, 3)

    # The other policy for each leg: mixed with the reference's forcing (its
    # solution is only as close to the FP64 reference's as the tolerance
    # makes it: both stop ~3e-10 from the converged solution, on different
    # Krylov paths) and fp64 with the safeguard (the like-for-like
    # denominator of the value).
    other = {}
    if not args.no_safeguard and not args.reference_forcing:
        for name, pp in (("mixed", params), ("fp64", psafe)):
            if name not in legs:
                continue
            s2, x2, r2, t2 = time_leg(H, ctx, args.n, args.lam, cfgs[name], pp, args.steps,
                                      args.warmup, graphs=args.graphs and name != "fp64")
            other[name] = {"s": round(float(np.mean(t2)), 6), "newton_iters": int(r2.iters),
                           "gmres_iters": int(r2.n_lin_iters),
                           "final_rel_residual": r2.res / r2.res0, "x": x2.numpy()}
            s2.close()
            x2.free()
        same = {"params": {"tight_res": SAFE_RES, "tight_eta": SAFE_ETA}}
        if "mixed" in other:
            same["mixed_reference_forcing"] = other["mixed"]
            if "fp64" in legs:
                same["mixed_reference_forcing"]["speedup_vs_fp64"] = round(
                    float(np.mean(legs["fp64"]["times"])) / other["mixed"]["s"], 3)
        if "fp64" in other:
            same["fp64_safeguard"] = other["fp64"]
            out["speedup_vs_fp64_same_policy"] = round(other["fp64"]["s"] / t_main, 3)
        out["same_policy"] = same

    if not args.no_cpu:
        info, path = run_cpu_baseline(args)
        xc = np.load(path)
        os.remove(path)
        parity = {"newton_iters_cpu_fp64": info["result"]["iters"],
                  "gmres_iters_cpu_fp64": info["result"]["n_lin_iters"]}
        xs_cpu = None
        if not args.reference_forcing:
            info_s, path_s = run_cpu_baseline(args, safeguard=True)
            xs_cpu = np.load(path_s)
            os.remove(path_s)
            parity["cpu_fp64_safeguard"] = {"s": round(info_s["seconds"], 3),
                                            "newton_iters": info_s["result"]["iters"],
                                            "gmres_iters": info_s["result"]["n_lin_iters"],
                                            "final_rel_residual": info_s["result"]["res"] /
                                            info_s["result"]["res0"]}
            args._safe_solution = xs_cpu
        def rl2(a, b):
            return float(np.linalg.norm(a - b) / np.linalg.norm(b))
        # every GPU solution against the CPU FP64 solve with the same policy
        for name, leg in legs.items():
            xg = leg["x"].numpy()
            ref = xs_cpu if (leg_params[name] is psafe and xs_cpu is not None) else xc
            parity[f"{name}_rel_l2_vs_cpu_fp64"] = rl2(xg, ref)
            parity[f"{name}_newton_iters"] = int(leg["res"].iters)
            parity[f"{name}_policy"] = "safeguard" if ref is xs_cpu else "reference forcing"
        for name, key, ref in (("mixed", "mixed_reference_forcing", xc),
                               ("fp64", "fp64_safeguard", xs_cpu)):
            if name in other and ref is not None:
                out["same_policy"][key]["rel_l2_vs_cpu_fp64"] = rl2(other[name]["x"], ref)
        if "mixed" in legs and xs_cpu is not None:
            parity["mixed_rel_l2_vs_cpu_fp64_reference_forcing"] = rl2(legs["mixed"]["x"].numpy(), xc)
            parity["cpu_safeguard_vs_cpu_reference_forcing"] = rl2(xs_cpu, xc)
        out["parity"] = parity
        out["cpu_baseline"] = {"value": round(info["seconds"], 3), "unit": "s",
                               "cores": info["threads"], "kind": "port",
                               "sample": f"one full FP64 solve of the same workload "
                                         f"(n={args.n}, same GMG, {args.forcing} forcing) by "
                                         f"the CPU oracle with OpenMP",
                               "cpu_model": cpu_model(), "nproc": os.cpu_count(),
                               "newton_iters": info["result"]["iters"],
                               "gmres_iters": info["result"]["n_lin_iters"]}
    for v in other.values():
        v.pop("x", None)
    if not args.no_chebyshev:
        out["chebyshev_outer"] = chebyshev_outer_leg(H, ctx, args, params, psafe)
    if not args.no_chandra:
        ch = chandra_leg(H, ctx, args)
        if not args.no_cpu:
            threads = args.cpu_threads or min(16, os.cpu_count() or 1)
            env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
                       OMP_PLACES="cores")
            pr = subprocess.run([sys.executable, os.path.abspath(__file__), "--chandra-cpu-leg"],
                                env=env, capture_output=True, text=True, timeout=args.cpu_timeout)
            if pr.returncode == 0:
                info = json.loads(pr.stdout.strip().splitlines()[-1])
                per = info["seconds"] / info["instances"]
                ch["cpu_baseline"] = {"value": round(per * CH_B, 3), "unit": "s", "cores":
                                      info["threads"], "kind": "port",
                                      "sample": f"FP64 oracle Newton-GMRES(20) for {info['instances']} "
                                                f"of the {CH_B} c values ({info['seconds']:.2f} s), "
                                                f"scaled by {CH_B}/{info['instances']}"}
        out["chandrasekhar"] = ch
    if not args.no_config0:
        out["config0"] = config0_leg(H, ctx, args)
    if not args.no_sweep:
        out["sweep"] = sweep_leg(H, ctx, args, params)
    if not args.no_learned:
        out["learned"] = learned_leg(H, ctx, args, params)
    print(json.dumps(out))
    for leg in legs.values():
        leg["solver"].close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
