"""B1 coarse start on the device (materialize_coarse_start.cpp:60-167 and
the B1 workflow): prolong3d bit for bit against the oracle, and a coarse
solution at lambda = 1.875 prolonged 15 -> 31 and corrected by Newton-Krylov
reaches the reference's frozen B1 start, whose norm the reference recorded as
9.0522684695809268 (experiments/c1/starts/frozen/B1_start_hashes.csv:2)."""
import numpy as np
import pytest

import numa_hip as H

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nc", [1, 3, 7, 15])
def test_prolong3d_matches_oracle(ctx, nc):
    import oracle as O
    c = np.random.default_rng(nc).normal(size=nc ** 3)
    nf = 2 * nc + 1
    dc, df = ctx.array(c), ctx.empty(nf ** 3)
    H.prolong3d(ctx, dc, nc, df)
    assert np.array_equal(df.numpy(), O.prolong3d(c, nc))


def test_b1_coarse_start_reaches_reference_start(ctx, tmp_path):
    cfg = H.solver_config(precond=H.PRECOND_GMG, jv_mode=H.DERIV_ANALYTIC, mg=H.mg_reference())
    cp = H.continuation_params(lambda_stop=1.875, step=0.125)
    xc = ctx.empty(15 ** 3)
    pts, status, _ = H.continuation(ctx, 15, cfg, cp, x_out=xc)
    assert status == 0 and abs(pts[-1].lambda_ - 1.875) < 1e-12
    xf = ctx.empty(31 ** 3)
    H.prolong3d(ctx, xc, 15, xf)
    start = xf.numpy()
    path = str(tmp_path / "B1_n31_coarse_start.txt")
    H.write_start(path, start)
    x = H.read_start(path, 31 ** 3)
    assert np.array_equal(x, start)
    x, r = H.newton_krylov(ctx, 31, 1.875, x, H.corrector_params(), cfg, restart=100)
    assert r.converged
    assert abs(np.linalg.norm(x) - 9.0522684695809268) <= 1e-9 * 9.0522684695809268
