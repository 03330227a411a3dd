"""GPU multigrid V-cycle against the oracle's V-cycle on the same right-hand
side: the reference PoissonMG (RB-GS, odd-n hierarchy, poisson_mg.hpp) and the
even-n weighted-Jacobi / Chebyshev hierarchies.  FP64 storage: relative 1e-12;
FP32 storage: relative 2e-5 (single-precision rounding through the cycle)."""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu


def cfgs():
    import numa_hip as H
    return {
        "reference": H.mg_reference(),
        "wjacobi": H.mg_wjacobi(),
        "cheb3": H.mg_cheb(levels=5),
        "cheb3_full": H.mg_cheb(levels=0),
        "cheb3_exact": H.mg_cheb(levels=5, coarse_iters=-1),
        "wj_exact": H.mg_wjacobi(levels=3, coarse_iters=-1),
    }


CASES = [("reference", 7), ("reference", 15), ("reference", 16), ("reference", 31),
         ("wjacobi", 16), ("wjacobi", 32), ("wjacobi", 33), ("cheb3", 16), ("cheb3", 32),
         ("cheb3", 64), ("cheb3", 48), ("cheb3_full", 32), ("cheb3_exact", 64),
         ("cheb3_exact", 256), ("wj_exact", 40)]


def _to_oracle(cfg):
    import oracle as O
    return O.MGConfig(**{k: getattr(cfg, k) for k, _ in cfg._fields_})


@pytest.mark.parametrize("name,n", CASES)
def test_vcycle_fp64(ctx, oracle, name, n):
    import numa_hip as H
    import oracle as O
    cfg = cfgs()[name]
    r = O.sin_point(n ** 3, 1.0)
    mg = H.PoissonMG(ctx, n, cfg, H.PRECISION_FP64)
    assert mg.sizes() == oracle.mg_levels(n, _to_oracle(cfg))
    dz = ctx.empty(n ** 3)
    mg(ctx.array(r), dz)
    ref = oracle.mg_apply(n, r, _to_oracle(cfg))
    assert rel(dz.numpy(), ref) < 1e-12


@pytest.mark.parametrize("name,n", [("wjacobi", 32), ("cheb3", 32), ("cheb3", 64),
                                    ("cheb3_exact", 64)])
def test_vcycle_fp32(ctx, oracle, name, n):
    import numa_hip as H
    import oracle as O
    cfg = cfgs()[name]
    r = O.sin_point(n ** 3, 1.0)
    mg = H.PoissonMG(ctx, n, cfg, H.PRECISION_MIXED)
    dz = ctx.empty(n ** 3, np.float32)
    mg(ctx.array(r.astype(np.float32)), dz)
    ref = oracle.mg_apply(n, r.astype(np.float32).astype(np.float64), _to_oracle(cfg))
    assert rel(dz.numpy(), ref) < 2e-5


def test_levels_reference_rule(ctx):
    import numa_hip as H
    assert H.PoissonMG(ctx, 15).sizes() == [15, 7, 3, 1]  # test_operators.cpp:193-194
    assert H.PoissonMG(ctx, 256, H.mg_cheb(levels=5)).sizes() == [256, 128, 64, 32, 16]


@pytest.mark.parametrize("n", [1, 2, 7, 16])
@pytest.mark.parametrize("precision", [0, 1])
def test_exact_coarse_solve_inverts_the_operator(ctx, oracle, n, precision):
    # one level with coarse_iters < 0: the V-cycle is the sine-transform solve,
    # so A z = r to rounding
    import numa_hip as H
    dt = np.float64 if precision == H.PRECISION_FP64 else np.float32
    cfg = H.mg_cheb(levels=1, coarse_iters=-1)
    r = np.random.default_rng(n).normal(size=n ** 3)
    mg = H.PoissonMG(ctx, n, cfg, precision)
    dz = ctx.empty(n ** 3, dt)
    mg(ctx.array(r.astype(dt)), dz)
    z = dz.numpy().astype(np.float64)
    assert rel(oracle.poisson_apply(n, z), r.astype(dt).astype(np.float64)) < (
        1e-13 if precision == H.PRECISION_FP64 else 1e-6)


def test_exact_coarse_solve_rejects_large_coarsest_level(ctx):
    import numa_hip as H
    with pytest.raises(H.NumaError):
        H.PoissonMG(ctx, 64, H.mg_cheb(levels=2, coarse_iters=-1))
