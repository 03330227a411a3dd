"""Temporally blocked smoothing (one launch per sweep or per pre/post phase)
against one launch per step: the V-cycle output is bit for bit the same for
every fusion mode, plane chunk, precision, smoother and degree, on grid sizes
that leave partial tiles, chunks and boxes; the autotuned choice is recorded in the
JSON tune cache; a solve with autotuned fusion takes the unfused solve's
iterations to the same solution (its FD steps take ||z|| from the fused output
launch)."""
import json
import os

import numpy as np
import pytest

import numa_hip as H

pytestmark = pytest.mark.gpu


def _vcycle(ctx, n, cfg, precision, r):
    dt = np.float32 if precision == H.PRECISION_MIXED else np.float64
    mg = H.PoissonMG(ctx, n, cfg, precision)
    dz = ctx.empty(n ** 3, dt)
    mg(ctx.array(r.astype(dt)), dz)
    out = dz.numpy()
    fus = mg.fusion()
    mg.close()
    return out, fus


CASES = [
    # (smoother factory kwargs, n)
    (dict(kind="cheb", degree=3, pre=2, post=2), 16),
    (dict(kind="cheb", degree=3, pre=2, post=2), 45),
    (dict(kind="cheb", degree=3, pre=2, post=2), 64),
    (dict(kind="cheb", degree=3, pre=2, post=2), 130),
    (dict(kind="cheb", degree=2, pre=1, post=3), 40),
    (dict(kind="cheb", degree=4, pre=1, post=1), 50),
    (dict(kind="cheb", degree=6, pre=1, post=1), 36),
    (dict(kind="wj", pre=2, post=2), 48),
    (dict(kind="wj", pre=3, post=1), 33),
]


def _cfg(spec, fuse, ichunk, rows=0):
    if spec["kind"] == "cheb":
        return H.mg_cheb(levels=4, degree=spec["degree"], pre=spec["pre"], post=spec["post"],
                         fuse=fuse, ichunk=ichunk, rows=rows)
    return H.mg_wjacobi(levels=4, pre=spec["pre"], post=spec["post"], fuse=fuse, ichunk=ichunk,
                        rows=rows)


@pytest.mark.parametrize("precision", [H.PRECISION_MIXED, H.PRECISION_FP64])
@pytest.mark.parametrize("spec,n", CASES)
def test_fused_smoothing_is_bitwise_unfused(ctx, spec, n, precision):
    rng = np.random.default_rng(n)
    r = rng.normal(size=n ** 3)
    ref, fus = _vcycle(ctx, n, _cfg(spec, H.FUSE_NONE, 0), precision, r)
    assert all(m == H.FUSE_NONE for m, _, _ in fus)
    r0 = 2 if precision == H.PRECISION_MIXED else 1
    per = spec.get("degree", 1)
    for fuse in (H.FUSE_SWEEP, H.FUSE_PHASE):
        group = per if fuse == H.FUSE_SWEEP else (6 // per) * per
        for ichunk, rows in ((16, r0), (32, r0), (64, 2 * r0), (7, 2 * r0), (32, 2 * r0)):
            if rows == 2 * r0 and group > 4:  # more than 160 KB of LDS: rejected
                with pytest.raises(H.NumaError):
                    _vcycle(ctx, n, _cfg(spec, fuse, ichunk, rows), precision, r)
                continue
            out, fus = _vcycle(ctx, n, _cfg(spec, fuse, ichunk, rows), precision, r)
            assert fus[0] == (fuse, ichunk, rows)
            np.testing.assert_array_equal(out, ref,
                                          err_msg=f"fuse={fuse} ichunk={ichunk} rows={rows}")
        # box form: one workgroup per box of output nodes, up to 6 steps
        for box, (ei, ej, ek) in H.BOX_EDGES.items():
            nbytes = 8 if precision == H.PRECISION_FP64 else 4
            nodes = (ei + 2 * group) * (ej + 2 * group) * (ek + 2 * group)
            fits = group <= 6 and 2 * nodes * nbytes <= 160 * 1024  # two LDS copies
            if not fits:
                with pytest.raises(H.NumaError):
                    _vcycle(ctx, n, _cfg(spec, fuse, 0, box), precision, r)
                continue
            out, fus = _vcycle(ctx, n, _cfg(spec, fuse, 0, box), precision, r)
            assert fus[0] == (fuse, 0, box)
            np.testing.assert_array_equal(out, ref, err_msg=f"fuse={fuse} box={box}")


def test_autotune_records_choice(ctx):
    path = os.environ["NUMA_HIP_TUNE_CACHE"]
    n = 96
    r = np.random.default_rng(1).normal(size=n ** 3)
    ref, _ = _vcycle(ctx, n, H.mg_cheb(levels=4), H.PRECISION_MIXED, r)
    out, fus = _vcycle(ctx, n, H.mg_cheb(levels=4, fuse=H.FUSE_AUTO), H.PRECISION_MIXED, r)
    np.testing.assert_array_equal(out, ref)
    assert all(m in (H.FUSE_NONE, H.FUSE_SWEEP, H.FUSE_PHASE) for m, _, _ in fus), fus
    with open(path) as f:
        cache = json.load(f)
    keys = [k for k in cache if "|mg-smooth-v5|f32|n=96|" in k]
    assert keys, cache
    mode, ichunk, rows = (int(v) for v in cache[keys[0]].split(","))
    assert (mode, ichunk, rows) == fus[0]
    # a new hierarchy of the same shape takes the cached choice
    out2, fus2 = _vcycle(ctx, n, H.mg_cheb(levels=4, fuse=H.FUSE_AUTO), H.PRECISION_MIXED, r)
    assert fus2 == fus
    np.testing.assert_array_equal(out2, ref)


@pytest.mark.parametrize("precision", [H.PRECISION_MIXED, H.PRECISION_FP64])
def test_solve_with_autotuned_fusion_matches_unfused(ctx, precision):
    n, lam = 48, 6.0
    p = H.krylov_params(gmres_restart=30)
    sols, results = [], []
    for fuse in (H.FUSE_NONE, H.FUSE_AUTO):
        cfg = H.solver_config(precond=H.PRECOND_GMG, precision=precision,
                              flexible=1 if precision == H.PRECISION_MIXED else 0,
                              mg=H.mg_cheb(levels=4, fuse=fuse))
        s = H.BratuSolver(ctx, n, lam, cfg, restart=30)
        x = np.zeros(n ** 3)
        res = s.solve_host(p, x)
        sols.append(x)
        results.append((res.iters, res.n_lin_iters, res.n_JVP, res.res))
        if fuse == H.FUSE_AUTO:
            assert all(m != H.FUSE_AUTO for m, _, _ in s.mg_fusion())
        s.close()
    # same V-cycles bit for bit; ||z|| of the FD step comes from the fused
    # output launch instead of a separate reduction (another summation order),
    # so the iterates differ in the last bits: same counts, and the converged
    # ||F|| (set by the last inexact linear solve) and solutions agree closely
    assert results[0][:3] == results[1][:3]
    assert abs(results[0][3] - results[1][3]) <= 1e-2 * results[0][3]
    assert np.linalg.norm(sols[0] - sols[1]) <= 1e-10 * np.linalg.norm(sols[0])


@pytest.mark.parametrize("precision", [H.PRECISION_MIXED, H.PRECISION_FP64])
@pytest.mark.parametrize("fuse,ichunk,rows", [(H.FUSE_NONE, 0, 0), (H.FUSE_SWEEP, 0, 0),
                                              (H.FUSE_PHASE, 16, 0), (H.FUSE_SWEEP, 0, -2)])
def test_vcycle_norm_from_the_output_launch(ctx, precision, fuse, ichunk, rows):
    # apply_norm: the same z as apply, and ||z|| of the stored values (taken
    # by the fused output launch, or by a separate reduction without fusion)
    n = 45
    dt = np.float32 if precision == H.PRECISION_MIXED else np.float64
    r = np.random.default_rng(3).normal(size=n ** 3).astype(dt)
    cfg = H.mg_cheb(levels=4, fuse=fuse, ichunk=ichunk, rows=rows)
    mg = H.PoissonMG(ctx, n, cfg, precision)
    dr, z1, z2 = ctx.array(r), ctx.empty(n ** 3, dt), ctx.empty(n ** 3, dt)
    mg(dr, z1)
    for rep in range(2):
        nz = mg.apply_norm(dr, z2)
        z = z2.numpy()
        np.testing.assert_array_equal(z, z1.numpy())
        ref = np.sqrt(np.sum(z.astype(np.float64) ** 2))
        assert abs(nz - ref) <= 1e-13 * ref, (rep, nz, ref)
    mg.close()
