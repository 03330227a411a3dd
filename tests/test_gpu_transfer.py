"""Residual and restriction in one launch (k_resid_restrict) against the
separate residual and restriction launches: the V-cycle output is bit for bit
the same for every plane chunk, precision and smoother (fused smoothing
included), on odd (reference coarsening, coarse I <-> fine 2I+1 with a last
plane 2I+2 < n) and even grids (the last coarse node has no plane 2I+2) whose
coarse levels leave partial 32 x 16 tiles and chunks.  The V-cycle against
the CPU oracle is test_gpu_mg."""
import numpy as np
import pytest

import numa_hip as H

pytestmark = pytest.mark.gpu


def _vcycle(ctx, n, cfg, precision, r, ich):
    dt = np.float32 if precision == H.PRECISION_MIXED else np.float64
    mg = H.PoissonMG(ctx, n, cfg, precision)
    mg.set_resid_restrict(ich)
    dz = ctx.empty(n ** 3, dt)
    mg(ctx.array(r.astype(dt)), dz)
    out = dz.numpy()
    mg.close()
    return out


CFGS = {
    "ref-rbgs": lambda: H.mg_reference(),
    "cheb3": lambda: H.mg_cheb(levels=5),
    "wj": lambda: H.mg_wjacobi(levels=4),
    "cheb3-fused": lambda: H.mg_cheb(levels=4, fuse=H.FUSE_SWEEP),
    "cheb3-phase": lambda: H.mg_cheb(levels=4, fuse=H.FUSE_PHASE, ichunk=16),
    "cheb2-13": lambda: H.mg_cheb(levels=3, degree=2, pre=1, post=3, fuse=H.FUSE_PHASE),
    "wj-fused": lambda: H.mg_wjacobi(levels=4, fuse=H.FUSE_SWEEP, ichunk=7),
}


@pytest.mark.parametrize("precision", [H.PRECISION_MIXED, H.PRECISION_FP64])
@pytest.mark.parametrize("name,n", [("ref-rbgs", 15), ("ref-rbgs", 63), ("cheb3", 16),
                                    ("cheb3", 70), ("cheb3", 130), ("wj", 45), ("wj", 97),
                                    ("cheb3-fused", 66), ("cheb3-fused", 45),
                                    ("cheb3-phase", 64), ("cheb3-phase", 99), ("cheb2-13", 50),
                                    ("wj-fused", 37)])
def test_fused_resid_restrict_is_bitwise(ctx, name, n, precision):
    rng = np.random.default_rng(1000 + n)
    r = rng.normal(size=n ** 3)
    cfg = CFGS[name]()
    ref = _vcycle(ctx, n, cfg, precision, r, 0)
    assert np.all(np.isfinite(ref))
    for ich in (-1, 1, 3, 8, 64):
        out = _vcycle(ctx, n, cfg, precision, r, ich)
        np.testing.assert_array_equal(out, ref, err_msg=f"ich={ich}")


def test_resid_restrict_rejects_huge_chunk(ctx):
    mg = H.PoissonMG(ctx, 16, H.mg_cheb(levels=3))
    with pytest.raises(H.NumaError):
        mg.set_resid_restrict(5000)
    mg.close()
