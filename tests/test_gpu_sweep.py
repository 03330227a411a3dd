"""hipGraph replay of the Arnoldi steps and the batched sweep on the GPU.

Graph replays are checked against the eager path bit for bit, over several
solves in a row with other work between them; a sweep over ragged instances
with several concurrent workers gives every instance the result of the same
solver run alone; small instances agree with the CPU oracle."""
import numpy as np
import pytest

from conftest import rel
import numa_hip as H

pytestmark = pytest.mark.gpu


def _cfg(precision, fuse=H.FUSE_AUTO, coarse_iters=24):
    mixed = precision == H.PRECISION_MIXED
    return H.solver_config(precond=H.PRECOND_GMG, precision=precision, flexible=1 if mixed else 0,
                           mg=H.mg_cheb(levels=4, fuse=fuse, coarse_iters=coarse_iters))


@pytest.mark.parametrize("precision", [H.PRECISION_MIXED, H.PRECISION_FP64])
@pytest.mark.parametrize("method", [H.METHOD_NEWTON, H.METHOD_CHEBYSHEV])
def test_graph_replay_is_bitwise_eager(ctx, precision, method):
    n, lam = 40, 6.0
    p = H.krylov_params(gmres_restart=30)
    cfg = _cfg(precision)
    cfg.method = method
    x0 = H.uniform_host(11, n ** 3, 0.1)
    eager = H.BratuSolver(ctx, n, lam, cfg, restart=30)
    xe = x0.copy()
    re_ = eager.solve_host(p, xe)
    g = H.BratuSolver(ctx, n, lam, cfg, restart=30)
    g.set_graphs(True)
    for rep in range(3):
        xg = x0.copy()
        rg = g.solve_host(p, xg)
        np.testing.assert_array_equal(xg, xe, err_msg=f"replay {rep}")
        assert (rg.iters, rg.n_lin_iters, rg.n_JVP, rg.n_F, rg.n_D2, rg.n_prec, rg.res) == \
            (re_.iters, re_.n_lin_iters, re_.n_JVP, re_.n_F, re_.n_D2, re_.n_prec, re_.res)
        # eager work on the same stream between replays
        xo = x0.copy()
        eager.solve_host(p, xo)
    assert g.graphs_cached() > 0
    eager.close()
    g.close()


def _instances(count, sizes, seed):
    n, lam = H.sweep_instances(seed, count, sizes=sizes)
    x0 = [H.uniform_host(seed * 7919 + 13 + i, int(n[i]) ** 3, 0.1) for i in range(count)]
    return n, lam, x0


# coarse_iters -1: the exact coarse solve, whose set-up (a table upload) runs
# in one worker while others capture their Arnoldi steps
@pytest.mark.parametrize("graphs,coarse", [(True, 24), (False, 24), (True, -1)])
def test_sweep_matches_solves_alone(ctx, graphs, coarse):
    n, lam, x0 = _instances(12, (16, 24, 32), 7)
    p = H.krylov_params(gmres_restart=30)
    cfg = _cfg(H.PRECISION_MIXED, coarse_iters=coarse)
    sw = H.Sweep(ctx, cfg, workers=4, restart=30)
    xs = [ctx.array(x) for x in x0]
    for run in range(2):  # the second run reuses solvers and captured graphs
        for x, v in zip(xs, x0):
            x.upload(v)
        res, sec = sw.run(n, lam, xs, p, graphs=graphs)
        assert sec > 0
        for i in range(len(xs)):
            s = H.BratuSolver(ctx, int(n[i]), float(lam[i]), cfg, restart=30)
            xa = x0[i].copy()
            ra = s.solve_host(p, xa)
            s.close()
            np.testing.assert_array_equal(xs[i].numpy(), xa, err_msg=f"run {run} instance {i}")
            assert (res[i].iters, res[i].n_lin_iters, res[i].res) == (ra.iters, ra.n_lin_iters, ra.res)
            assert res[i].converged
    sw.close()


def test_sweep_instances_match_cpu_oracle(ctx, oracle):
    import oracle as O
    n, lam, x0 = _instances(3, (12, 16), 3)
    p = H.krylov_params(gmres_restart=30)
    cfg = _cfg(H.PRECISION_FP64, fuse=H.FUSE_NONE)
    sw = H.Sweep(ctx, cfg, workers=2, restart=30)
    xs = [ctx.array(x) for x in x0]
    res, _ = sw.run(n, lam, xs, p, graphs=True)
    sw.close()
    mg = O.MGConfig(**{k: getattr(cfg.mg, k) for k, _ in cfg.mg._fields_})
    ocfg = O.SolverConfig(method=0, precond=2, precision=0, flexible=0, jv_mode=0, d2_mode=0,
                          mg=mg)
    for i in range(len(xs)):
        op = O.krylov_params(gmres_restart=30)
        xc, rc = oracle.bratu_solve(int(n[i]), float(lam[i]), ocfg, op, x0[i])
        assert res[i].iters == rc.iters, (i, res[i].iters, rc.iters)
        assert res[i].n_lin_iters == rc.n_lin_iters
        assert rel(xs[i].numpy(), xc) < 1e-9
