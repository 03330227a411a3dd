"""GPU batched Chandrasekhar H-equation against the CPU oracle.

FP64 (MFMA f64 contraction): operator values to 1e-13 relative, and each
instance's newton_krylov counters identical to the oracle's solve for that c,
iterates to 1e-11.  Mixed (split-BF16 contraction with FP32 accumulation for
J v only, F stays FP64): J v within 3e-5 relative of FP64 (three bf16 products
keep ~16 bits of K and x), solves converge to the same tolerance with the
Newton count within +-1 and the solution within 1e-9 of the FP64 oracle.
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, rel

pytestmark = pytest.mark.gpu

KEYS = ("converged", "iters", "n_F", "n_JVP", "n_D2", "n_lin_iters", "n_lin_failures", "n_prec")


def _oparams(p):
    import oracle as O
    return O.KrylovParams(**{k: getattr(p, k) for k, _ in p._fields_})


def _ocfg():
    import oracle as O
    return O.solver_config(method=0, precond=O.PRECOND_NONE, jv_mode=O.DERIV_ANALYTIC)


def _x(b, n, seed=0):
    rng = np.random.default_rng(seed)
    return 1.0 + 0.2 * rng.standard_normal((b, n))


@pytest.mark.parametrize("n,rule,form", [(8, 0, 0), (24, 0, 0), (100, 1, 1), (300, 1, 0),
                                         (257, 0, 1)])
def test_eval_jvp_fp64(ctx, oracle, n, rule, form):
    import numa_hip as H
    cs = np.array([0.5, 0.8, 0.99, 0.3, 0.9999])
    ch = H.ChandrasekharBatch(ctx, n, cs, rule=rule, form=form)
    x = _x(cs.size, n)
    v = _x(cs.size, n, seed=1) - 1.0
    F, norms = ch.eval(x)
    J = ch.jvp(x, v)
    for b, c in enumerate(cs):
        Fo = oracle.chandra_eval(n, c, x[b], rule=rule, form=form)
        assert rel(F[b], Fo) < 1e-13
        assert abs(norms[b] - np.linalg.norm(Fo)) <= 1e-12 * np.linalg.norm(Fo)
        assert rel(J[b], oracle.chandra_jvp(n, c, x[b], v[b], rule=rule, form=form)) < 1e-13


def test_jvp_split_bf16_error(ctx, oracle):
    import numa_hip as H
    n = 512
    cs = np.linspace(0.5, 0.9999, 64)
    x = _x(cs.size, n)
    v = _x(cs.size, n, seed=2) - 1.0
    ch = H.ChandrasekharBatch(ctx, n, cs, rule=1, form=1, precision=H.PRECISION_MIXED)
    J = ch.jvp(x, v)
    for b in (0, 17, 63):
        ref = oracle.chandra_jvp(n, cs[b], x[b], v[b], rule=1, form=1)
        assert rel(J[b], ref) < 3e-5


def test_golden_n8(ctx):
    """tests/test_chandrasekhar_op.cpp:101-131: N = 8, c = 0.5, Gauss-Legendre."""
    import numa_hip as H
    gold = np.array(json.load(open(os.path.join(GOLDEN, "reference_data.json")))["chandra_n8_c05"])
    ch = H.ChandrasekharBatch(ctx, 8, [0.5, 0.5], rule=0, form=0)
    F, norms = ch.eval(np.stack([gold, gold]))
    assert norms[0] < 2e-10
    x, res = ch.solve(H.krylov_params())
    assert res[0].converged and rel(x[0], gold) < 2e-7


@pytest.mark.parametrize("n,rule,form,restart", [(8, 0, 0, 50), (64, 1, 1, 20), (200, 1, 1, 20),
                                                 (128, 0, 0, 5)])
def test_batched_solve_fp64_matches_oracle(ctx, oracle, n, rule, form, restart):
    import numa_hip as H
    cs = np.array([0.5, 0.75, 0.9, 0.99, 0.9999, 0.2])
    p = H.krylov_params(rtol_nl=1e-10, gmres_restart=restart)
    ch = H.ChandrasekharBatch(ctx, n, cs, rule=rule, form=form, restart=restart)
    x, res = ch.solve(p)
    for b, c in enumerate(cs):
        xo, ro = oracle.chandra_solve(n, c, _ocfg(), _oparams(p), rule=rule, form=form)
        got = {k: getattr(res[b], k) for k in KEYS}
        want = {k: getattr(ro, k) for k in KEYS}
        assert got == want, (c, got, want)
        assert rel(x[b], xo) < 1e-11
        # moment identity sum_i w_i x_i = (2/c)(1 - sqrt(1 - c)) at the solution
        if rule == 1 and res[b].converged:
            assert abs(np.mean(x[b]) - 2 / c * (1 - np.sqrt(1 - c))) < 1e-8


def test_batched_solve_mixed(ctx, oracle):
    import numa_hip as H
    n = 1024
    cs = np.array([0.5, 0.8, 0.95, 0.99, 0.999, 0.9999])
    p = H.krylov_params(rtol_nl=1e-10, gmres_restart=20)
    ch = H.ChandrasekharBatch(ctx, n, cs, rule=1, form=1, precision=H.PRECISION_MIXED, restart=20)
    x, res = ch.solve(p)
    for b, c in enumerate(cs):
        xo, ro = oracle.chandra_solve(n, c, _ocfg(), _oparams(p), rule=1, form=1)
        assert res[b].converged and res[b].res <= 1e-10 * res[b].res0
        assert abs(res[b].iters - ro.iters) <= 1
        assert rel(x[b], xo) < 1e-9


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("n,batch", [(700, 70), (1000, 40)])
def test_stored_k_matches_generated(ctx, monkeypatch, precision, n, batch):
    """K stored once per handle and K generated inside the contraction kernels
    feed the same operand values to the same MFMA sequence: F, J v and the
    batched solve agree bit for bit.  N = 700 pads to 768 and batch 70 to 128:
    4 column tiles per wave, split-K 4, three k chunks of 64 (odd count);
    N = 1000 pads to 1024 and batch 40 to 64: 2 column tiles, two chunks of 128."""
    import numa_hip as H
    prec = H.PRECISION_FP64 if precision == "fp64" else H.PRECISION_MIXED
    cs = np.linspace(0.3, 0.9999, batch)
    x = _x(cs.size, n, seed=3)
    v = _x(cs.size, n, seed=4) - 1.0
    p = H.krylov_params(rtol_nl=1e-10, gmres_restart=20)
    got = []
    for store in ("1", "0"):
        monkeypatch.setenv("NUMA_HIP_CHANDRA_STORE_K", store)
        ch = H.ChandrasekharBatch(ctx, n, cs, rule=1, form=1, precision=prec, restart=20)
        F, norms = ch.eval(x)
        J = ch.jvp(x, v)
        xs, res = ch.solve(p)
        got.append((F, norms, J, xs, [(r.iters, r.n_JVP, r.converged) for r in res]))
        ch.close()
    for a, b in zip(got[0], got[1]):
        if isinstance(a, list):
            assert a == b
        else:
            assert np.array_equal(a, b)
    assert all(c for _, _, c in got[0][4])


@pytest.mark.parametrize("method", ["newton", "chebyshev"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_packed_contractions_match_full(ctx, monkeypatch, precision, method):
    """Once some instances have converged, the lockstep solve contracts only
    the columns whose results are used (gathered into a packed block of whole
    column tiles, results scattered back) and runs the Arnoldi step's J v,
    dots, MGS passes and finishes only for the instances still in their
    cycle.  Every instance's iterate and counters are bit for bit those of
    the run that computes everything for every instance
    (NUMA_HIP_CHANDRA_PACK=0).  256 instances (padded batch 256: FP64 tiles of
    64 columns, split-BF16 tiles of 128) with c from 0.2 to 0.9999, whose
    Newton counts differ, so the packed width drops below the batch in both
    precisions; the Chebyshev method also packs F''[z, z] and its trial F."""
    import numa_hip as H
    prec = H.PRECISION_FP64 if precision == "fp64" else H.PRECISION_MIXED
    meth = H.METHOD_NEWTON if method == "newton" else H.METHOD_CHEBYSHEV
    n, batch = 512, 256
    cs = np.linspace(0.2, 0.9999, batch)
    p = H.krylov_params(rtol_nl=1e-10, gmres_restart=20)
    got = []
    for pack in ("1", "0"):
        monkeypatch.setenv("NUMA_HIP_CHANDRA_PACK", pack)
        print(f"chandrasekhar packed={pack} {precision} {method} n={n} batch={batch}", flush=True)
        ch = H.ChandrasekharBatch(ctx, n, cs, rule=1, form=1, precision=prec, restart=20,
                                  method=meth)
        xs, res = ch.solve(p)
        got.append((xs, [(r.iters, r.n_F, r.n_JVP, r.n_D2, r.n_lin_iters, r.converged, r.res)
                         for r in res]))
        ch.close()
    assert np.array_equal(got[0][0], got[1][0])
    assert got[0][1] == got[1][1]
    its = [r[0] for r in got[0][1]]
    assert all(r[5] for r in got[0][1])
    assert min(its) < max(its)  # instances drop out at different steps


@pytest.mark.parametrize("n,rule,form", [(8, 0, 0), (64, 1, 1), (200, 1, 0)])
def test_eval_d2_fp64(ctx, oracle, n, rule, form):
    import numa_hip as H
    cs = np.array([0.5, 0.8, 0.99, 0.9999])
    ch = H.ChandrasekharBatch(ctx, n, cs, rule=rule, form=form)
    x = _x(cs.size, n)
    v = _x(cs.size, n, seed=1) - 1.0
    D = ch.d2(x, v)
    for b, c in enumerate(cs):
        assert rel(D[b], oracle.chandra_d2(n, c, x[b], v[b], rule=rule, form=form)) < 1e-13


@pytest.mark.parametrize("n,rule,form,restart", [(8, 0, 0, 50), (64, 1, 1, 20), (200, 1, 0, 20)])
def test_batched_chebyshev_fp64_matches_oracle(ctx, oracle, n, rule, form, restart):
    """chebyshev_inexact (krylov_newton.hpp:116-189) in lockstep over the batch:
    every counter identical to the oracle's solve for that c (whose form 0 /
    Gauss-Legendre solves equal the reference build's; tests/test_oracle.py)."""
    import numa_hip as H
    import oracle as O
    cs = np.array([0.5, 0.75, 0.9, 0.99, 0.9999, 0.2])
    p = H.krylov_params(rtol_nl=1e-10, gmres_restart=restart)
    ch = H.ChandrasekharBatch(ctx, n, cs, rule=rule, form=form, restart=restart,
                              method=H.METHOD_CHEBYSHEV)
    x, res = ch.solve(p)
    cfg = O.solver_config(method=1, precond=O.PRECOND_NONE, jv_mode=O.DERIV_ANALYTIC,
                          d2_mode=O.DERIV_ANALYTIC)
    for b, c in enumerate(cs):
        xo, ro = oracle.chandra_solve(n, c, cfg, _oparams(p), rule=rule, form=form)
        got = {k: getattr(res[b], k) for k in KEYS}
        want = {k: getattr(ro, k) for k in KEYS}
        assert got == want, (c, got, want)
        assert res[b].n_D2 >= 1
        assert rel(x[b], xo) < 1e-11


def test_batched_solves_match_reference_build(ctx):
    """The reference's own newton_krylov / chebyshev_inexact on ChandrasekharOp
    (tests/golden/chandra_solves.json, from the reference build), one batch per
    (N, method) over c in {0.5, 0.9, 0.9999}: every counter identical."""
    import numa_hip as H
    cases = json.load(open(os.path.join(GOLDEN, "chandra_solves.json")))
    for n in sorted({c["n"] for c in cases}):
        for method in (0, 1):
            sel = [c for c in cases if c["n"] == n and c["method"] == method]
            cs = np.array([c["c"] for c in sel])
            ch = H.ChandrasekharBatch(ctx, n, cs, rule=0, form=0, restart=50, method=method)
            x, res = ch.solve(H.krylov_params(rtol_nl=1e-10))
            ch.close()
            for b, c in enumerate(sel):
                got = {k: int(getattr(res[b], k)) for k in KEYS}
                assert got == {k: int(c["result"][k]) for k in KEYS}, (n, method, c["c"])
                assert rel(x[b], c["x"]) < 1e-11
