"""Learned preconditioner on the GPU: the conv net's forward pass, loss and
gradients against the FP64 oracle (oracle/convnet_ref.py, with the device's
bfloat16 rounding points reproduced), training on solver-generated
(residual, V-cycle correction) pairs, FGMRES with the V-cycle fallback, and
what the trained network gains over the same network untrained and over no
preconditioner."""
import numpy as np
import pytest

from conftest import rel
import convnet_ref as cr
import numa_hip as nh

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,B", [(8, 1), (13, 2), (33, 1)])
def test_forward_matches_oracle(ctx, n, B):
    net = nh.ConvNet(ctx, seed=11)
    rng = np.random.default_rng(n)
    x = rng.normal(size=(B, n, n, n)).astype(np.float32)
    out = net.forward(n, x)
    Ws, bs = cr.unflatten(net.get_params())
    for b in range(B):
        ref, _ = cr.forward(x[b].astype(np.float64), Ws, bs, bf16_mode=True)
        # bf16 activations on both sides; the residual difference is FP32
        # accumulation order and the odd bf16 rounding flip
        assert rel(out[b], ref) < 3e-3, (n, b, rel(out[b], ref))
        raw, _ = cr.forward(x[b].astype(np.float64), Ws, bs, bf16_mode=False)
        assert rel(out[b], raw) < 5e-2
    net.close()


@pytest.mark.parametrize("n,B,loss_kind", [(8, 2, nh.LOSS_CORRECTION), (12, 1, nh.LOSS_CORRECTION),
                                            (8, 2, nh.LOSS_RESIDUAL), (11, 1, nh.LOSS_RESIDUAL)])
def test_loss_and_gradients_match_oracle(ctx, n, B, loss_kind):
    net = nh.ConvNet(ctx, seed=5)
    rng = np.random.default_rng(100 + n)
    x = rng.normal(size=(B, n, n, n)).astype(np.float32)
    y = rng.normal(size=(B, n, n, n)).astype(np.float32)
    loss, g = net.loss_grad(n, x, y, loss_kind)
    Ws, bs = cr.unflatten(net.get_params())
    kind = "correction" if loss_kind == nh.LOSS_CORRECTION else "residual"
    lref, dWs, dbs, _ = cr.loss_and_grads(x.astype(np.float64), y.astype(np.float64), Ws, bs,
                                          bf16_mode=True, loss_kind=kind)
    assert abs(loss - lref) <= 2e-3 * lref, (loss, lref)
    gW, gb = cr.unflatten(g.astype(np.float64))
    for l in range(6):
        assert rel(gW[l], dWs[l]) < 1e-2, (l, rel(gW[l], dWs[l]))
        assert rel(gb[l], dbs[l]) < 1e-2, (l, rel(gb[l], dbs[l]))
    net.close()


def test_apply_is_scaled_network(ctx):
    n = 24
    net = nh.ConvNet(ctx, seed=4)
    rng = np.random.default_rng(9)
    r = rng.normal(size=n ** 3).astype(np.float32)
    dz = ctx.empty(n ** 3, np.float32)
    net.apply(n, ctx.array(r), dz)
    z1 = dz.numpy()
    net.apply(n, ctx.array(2 * r), dz)
    z2 = dz.numpy()
    np.testing.assert_array_equal(z2, 2 * z1)  # positively homogeneous, bit for bit
    rho = float(np.sqrt(np.mean(r.astype(np.float64) ** 2)))
    x = (r * np.float32(1.0 / rho)).reshape(1, n, n, n)
    out = net.forward(n, x).ravel()
    assert rel(z1, out * (rho / (n + 1) ** 2)) < 1e-3
    net.close()


def test_train_step_reduces_loss(ctx):
    n, B = 16, 4
    net = nh.ConvNet(ctx, seed=2)
    mg = nh.PoissonMG(ctx, n, nh.mg_cheb(levels=3), nh.PRECISION_MIXED)
    rng = np.random.default_rng(7)
    r = rng.normal(size=(B, n ** 3)).astype(np.float32)
    z = np.empty_like(r)
    dzi = ctx.empty(n ** 3, np.float32)
    for b in range(B):
        mg(ctx.array(r[b]), dzi)
        z[b] = dzi.numpy()
    dr, dz = ctx.array(r.ravel()), ctx.array(z.ravel())
    losses = [net.train_step(n, dr, dz, B, 3e-3) for _ in range(60)]
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.7 * losses[0], (losses[0], losses[-1])
    mg.close()
    net.close()


@pytest.fixture(scope="module")
def trained(ctx):
    net = nh.ConvNet(ctx, seed=21)
    cfg = nh.net_train_config(pairs=2000, batch=4, sizes=(16, 24, 32), lr=2e-3, seed=5,
                              mg=nh.mg_cheb(levels=3), loss=nh.LOSS_RESIDUAL)
    st, hist = net.train_solver(cfg, hist_len=200)
    yield net, st, hist
    net.close()


def test_train_from_solver(trained):
    net, st, hist = trained
    assert st.pairs == 2000 and st.steps >= 500 and st.solves >= 2, (st.pairs, st.steps, st.solves)
    assert np.all(np.isfinite(hist))
    assert hist[-10:].mean() < hist[:10].mean(), (hist[:10].mean(), hist[-10:].mean())


def _solver(ctx, n, lam, precond, mg=None):
    cfg = nh.solver_config(precond=precond, precision=nh.PRECISION_MIXED, flexible=1,
                           mg=mg if mg is not None else nh.mg_cheb(levels=3))
    return nh.BratuSolver(ctx, n, lam, cfg, restart=30)


@pytest.mark.parametrize("n,lam", [(24, 6.0), (64, 6.5)])
def test_learned_fgmres_with_fallback(ctx, oracle, trained, n, lam):
    """The learned solve reaches the FP64 oracle's solution to 1e-10 (relative
    L2): both use the final-solve safeguard (forcing capped at 1e-6 once
    ||F|| <= 1e-6 ||F0||, converged only once the last step started there),
    so the final iterate does not depend on the path
    (network steps, fallback, FP32 basis) by which the solve got there."""
    import oracle as O
    net = trained[0]
    p = nh.krylov_params(gmres_restart=30, tight_res=1e-6, tight_eta=1e-6)
    mg = nh.mg_cheb(levels=3)
    s = _solver(ctx, n, lam, nh.PRECOND_LEARNED, mg=mg)
    s.set_net(net, 0.5)
    x = np.zeros(n ** 3)
    r = s.solve_host(p, x)
    assert r.converged and r.res <= 1e-8 * r.res0 * 1.0001
    st = s.learned_stats()
    assert st["net_applications"] + st["gmg_applications"] == r.n_prec
    assert st["net_applications"] >= 1
    assert st["fallbacks"] <= r.iters + 1
    s.close()
    ocfg = O.solver_config(precond=O.PRECOND_GMG,
                           mg=O.MGConfig(**{k: getattr(mg, k) for k, _ in mg._fields_}))
    xo, ro = oracle.bratu_solve(n, lam, ocfg,
                                O.KrylovParams(**{k: getattr(p, k) for k, _ in p._fields_}))
    assert ro.converged
    assert abs(r.iters - ro.iters) <= 1
    assert rel(x, xo) <= 1e-10, rel(x, xo)


def test_learned_requires_flexible(ctx):
    cfg = nh.solver_config(precond=nh.PRECOND_LEARNED, precision=nh.PRECISION_MIXED, flexible=0,
                           mg=nh.mg_cheb(levels=3))
    with pytest.raises(nh.NumaError):
        nh.BratuSolver(ctx, 16, 1.0, cfg, restart=10)


@pytest.mark.parametrize("rhs", ["random", "constant"])
def test_trained_network_gains(ctx, trained, rhs):
    """FGMRES residual after 8 steps at 32^3, J(0) z = b with b random or
    b = F(0) = -lambda: trained network < untrained network (same
    initialisation) and < no preconditioner."""
    n, lam, steps = 32, 6.0, 8
    rng = np.random.default_rng(31)
    b = ctx.array(rng.normal(size=n ** 3) if rhs == "random" else -lam * np.ones(n ** 3))
    x0 = ctx.array(np.zeros(n ** 3))
    z = ctx.empty(n ** 3)
    out = {}
    untrained = nh.ConvNet(ctx, seed=21)
    for name, net in (("trained", trained[0]), ("untrained", untrained)):
        s = _solver(ctx, n, lam, nh.PRECOND_LEARNED)
        s.set_net(net, 1.0)  # no fallback
        out[name] = s.gmres(x0, b, z, 1e-14, 30, steps).rel_res
        assert s.learned_stats()["gmg_applications"] == 0
        s.close()
    s = _solver(ctx, n, lam, nh.PRECOND_NONE)
    out["none"] = s.gmres(x0, b, z, 1e-14, 30, steps).rel_res
    s.close()
    s = _solver(ctx, n, lam, nh.PRECOND_GMG)
    out["gmg"] = s.gmres(x0, b, z, 1e-14, 30, steps).rel_res
    s.close()
    untrained.close()
    print("rel_res after", steps, "FGMRES steps,", rhs, "rhs:", out)
    assert out["trained"] < out["untrained"], out
    assert out["trained"] < out["none"], out
    # Against the method without the network (FGMRES + V-cycle): the network
    # alone does not reach the V-cycle's reduction (measured at 64^3: 3.1e-3
    # vs 2.8e-8 with a random right-hand side).  The ratio is recorded; this
    # comparison is an expected failure, not a pass.
    ratio = out["trained"] / out["gmg"]
    print("trained / gmg residual ratio:", ratio)
    if not out["trained"] <= out["gmg"]:
        pytest.xfail(f"network alone loses to the V-cycle: {out['trained']:.3g} vs "
                     f"{out['gmg']:.3g} after {steps} FGMRES steps (ratio {ratio:.3g})")


def test_committed_weights_gain(ctx):
    """The committed parameters (200k pairs on 16^3-32^3 grids) load through
    set_params and, as the preconditioner alone, leave a smaller FGMRES
    residual after 8 steps on J(0) z = b (random b, 64^3) than no
    preconditioner (measured 0.0068 vs 0.038) and than the 64^3-128^3 set
    trained at the same budget (0.149)."""
    import os
    wdir = os.path.join(os.path.dirname(nh.__file__), "weights")
    n, lam, steps = 64, 6.0, 8
    rng = np.random.default_rng(31)
    b = ctx.array(rng.normal(size=n ** 3))
    x0 = ctx.array(np.zeros(n ** 3))
    z = ctx.empty(n ** 3)
    out = {}
    for name in ("convnet_200k_n16-32", "convnet_200k_n64-128"):
        net = nh.ConvNet(ctx, seed=0)
        net.set_params(np.load(os.path.join(wdir, name + ".npy")))
        s = _solver(ctx, n, lam, nh.PRECOND_LEARNED)
        s.set_net(net, 1.0)
        out[name] = s.gmres(x0, b, z, 1e-14, 30, steps).rel_res
        s.close()
        net.close()
    s = _solver(ctx, n, lam, nh.PRECOND_NONE)
    out["none"] = s.gmres(x0, b, z, 1e-14, 30, steps).rel_res
    s.close()
    print("rel_res after", steps, "FGMRES steps at 64^3:", out)
    assert out["convnet_200k_n16-32"] < 0.5 * out["none"], out
    assert out["convnet_200k_n16-32"] < out["convnet_200k_n64-128"], out
