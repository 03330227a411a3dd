"""The batched solver (numa_hip_batch): many Bratu3D instances of different
grid size and lambda solved together, one launch per operation over the
instance table and one graph launch per lockstep Arnoldi step.

Every instance's solution and counters equal, bit for bit, those of the
single-instance solver with the same configuration and unfused smoothing
(the batched launches run the single-instance kernel bodies on each
instance's own block range); a batch of one equals the same instance inside
a batch of many; graph replays equal eager launches; and a sample agrees
with the FP64 CPU oracle (final-solve safeguard on both sides)."""
import numpy as np
import pytest

from conftest import rel
import numa_hip as H

pytestmark = pytest.mark.gpu

SAFEGUARD = dict(tight_res=1e-6, tight_eta=1e-6)


def _cfg(levels=3, fuse=H.FUSE_NONE):
    return H.solver_config(method=H.METHOD_NEWTON, precond=H.PRECOND_GMG,
                           precision=H.PRECISION_MIXED, flexible=1, jv_mode=H.DERIV_FD,
                           mg=H.mg_cheb(levels=levels, fuse=fuse, coarse_iters=-1))


def _instances(count, sizes, seed):
    n, lam = H.sweep_instances(seed, count, sizes=sizes)
    x0 = [H.uniform_host(seed * 7919 + 13 + i, int(n[i]) ** 3, 0.1) for i in range(count)]
    return n, lam, x0


def _solve_batch(ctx, n, lam, x0, cfg, p, graphs, reps=1, shapes=None):
    b = H.Batch(ctx, n, lam, cfg, restart=30)
    b.set_graphs(graphs)
    if shapes is not None:
        b.set_smoothing(shapes)
    xs = [ctx.array(v) for v in x0]
    for _ in range(reps):
        for x, v in zip(xs, x0):
            x.upload(v)
        res, sec = b.solve(xs, p)
    got = [x.numpy() for x in xs]
    st = b.stats()
    for x in xs:
        x.free()
    b.close()
    return got, res, st


def _alone(ctx, n, lam, x0, cfg, p):
    s = H.BratuSolver(ctx, int(n), float(lam), cfg, restart=30)
    x = x0.copy()
    r = s.solve_host(p, x)
    s.close()
    return x, r


def _counters(r):
    return (r.converged, r.iters, r.n_F, r.n_JVP, r.n_lin_iters, r.n_lin_failures, r.n_prec,
            r.res0, r.res)


# fuse AUTO: the batch's smoothing in the box form where its autotuner picks
# it, bit for bit the single-instance solver's one launch per step
@pytest.mark.parametrize("graphs,fuse", [(False, H.FUSE_NONE), (True, H.FUSE_NONE),
                                         (True, H.FUSE_AUTO)])
def test_batch_equals_solves_alone(ctx, graphs, fuse):
    n, lam, x0 = _instances(8, (16, 24, 32), 3)
    print(f"batch of {len(n)}: n={list(map(int, n))} graphs={graphs} fuse={fuse}", flush=True)
    p = H.krylov_params(gmres_restart=30)
    cfg = _cfg()
    got, res, st = _solve_batch(ctx, n, lam, x0, _cfg(fuse=fuse), p, graphs, reps=2)
    assert st["arnoldi_steps"] > 0
    if graphs:
        assert st["graphs"] > 0 and st["graph_launches"] > 0
    for i in range(len(n)):
        xa, ra = _alone(ctx, n[i], lam[i], x0[i], cfg, p)
        np.testing.assert_array_equal(got[i], xa, err_msg=f"instance {i} (n={int(n[i])})")
        assert _counters(res[i]) == _counters(ra), i
        assert res[i].converged


def test_batch_of_one_equals_member_of_batch(ctx):
    n, lam, x0 = _instances(6, (16, 32), 9)
    p = H.krylov_params(gmres_restart=30, **SAFEGUARD)
    cfg = _cfg()
    got, res, _ = _solve_batch(ctx, n, lam, x0, cfg, p, True)
    for i in (0, len(n) - 1):
        g1, r1, _ = _solve_batch(ctx, n[i:i + 1], lam[i:i + 1], x0[i:i + 1], cfg, p, True)
        np.testing.assert_array_equal(g1[0], got[i])
        assert _counters(r1[0]) == _counters(res[i])


def test_batch_restarts_and_iteration_caps(ctx):
    """Short cycles (restart 2) force restarted GMRES with the residual
    recomputed by J v; a cap of 3 iterations per linear solve forces
    unconverged linear solves: both equal the single-instance solver."""
    n, lam, x0 = _instances(5, (16, 24), 21)
    cfg = _cfg()
    for kw in (dict(gmres_restart=2), dict(gmres_restart=30, gmres_max_iters=3)):
        p = H.krylov_params(**kw)
        b = H.Batch(ctx, n, lam, cfg, restart=30)
        xs = [ctx.array(v) for v in x0]
        res, _ = b.solve(xs, p)
        got = [x.numpy() for x in xs]
        b.close()
        for x in xs:
            x.free()
        for i in range(len(n)):
            xa, ra = _alone(ctx, n[i], lam[i], x0[i], cfg, p)
            np.testing.assert_array_equal(got[i], xa, err_msg=f"{kw} instance {i}")
            assert _counters(res[i]) == _counters(ra), (kw, i)


def test_batch_rejects_unsupported(ctx):
    cfg = _cfg()
    with pytest.raises(Exception):
        H.Batch(ctx, [16, 200], [6.0, 6.0], cfg)  # n beyond the batched stencil form
    with pytest.raises(Exception):
        H.Batch(ctx, [16, 32], [6.0, 6.0], _cfg(levels=0))  # different level counts
    bad = _cfg()
    bad.precision = H.PRECISION_FP64
    with pytest.raises(Exception):
        H.Batch(ctx, [16], [6.0], bad)


@pytest.mark.parametrize("shape", [1, 2, 3, 4, 5, 6, 7])
def test_batch_box_smoothing_is_bitwise(ctx, shape):
    """Every box shape of the batched smoothing on every level equals the
    single-instance solver's one launch per smoothing step."""
    n, lam, x0 = _instances(5, (16, 24, 32), 17)
    p = H.krylov_params(gmres_restart=30)
    got, res, _ = _solve_batch(ctx, n, lam, x0, _cfg(fuse=H.FUSE_AUTO), p, True,
                               shapes=[shape, shape])
    for i in range(len(n)):
        xa, ra = _alone(ctx, n[i], lam[i], x0[i], _cfg(), p)
        np.testing.assert_array_equal(got[i], xa, err_msg=f"shape {shape} instance {i}")
        assert _counters(res[i]) == _counters(ra), i
