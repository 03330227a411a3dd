"""GPU parity of the Bratu3D kernels against the CPU oracle (same inputs).

Tolerances: FP64 kernels follow the reference's operation order with FMA
contraction off, so they differ from the oracle only through the device exp():
relative 1e-14 for F, J v, F'' and the Jacobi diagonal.  Forward-difference
quotients amplify an ulp of F by 1/h: relative 1e-7 for J v (FP64), 2e-6 for
the mixed form (v and out FP32), 1e-4 for the central F'' (the reference's own
analytic-vs-FD bound, test_operators.cpp:56-60).
"""
import numpy as np
import pytest

from conftest import rel
import numa_hip as H

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 7, 8, 16, 31, 33, 64, 65]


def _inputs(n, seed=0):
    import oracle as O
    N = n ** 3
    x = O.sin_point(N, 0.3)
    v = O.sin_point(N, 1.0)[::-1].copy()
    return x, v


@pytest.mark.parametrize("n", SIZES)
def test_eval(ctx, oracle, n):
    import numa_hip as H
    x, v = _inputs(n)
    lam = 6.5
    f = H.Bratu3D(ctx, n, lam)
    dx, dF = ctx.array(x), ctx.empty(n ** 3)
    nrm = f.eval(dx, dF)
    ref = oracle.bratu_eval(n, lam, x)
    assert rel(dF.numpy(), ref) < 1e-14
    assert abs(nrm - np.linalg.norm(ref)) <= 1e-13 * np.linalg.norm(ref)


@pytest.mark.parametrize("n", SIZES)
def test_jvp_d2_diag(ctx, oracle, n):
    import numa_hip as H
    x, v = _inputs(n)
    lam = 3.0
    f = H.Bratu3D(ctx, n, lam)
    dx, dv, do = ctx.array(x), ctx.array(v), ctx.empty(n ** 3)
    f.jvp(dx, dv, do)
    assert rel(do.numpy(), oracle.bratu_jvp(n, lam, x, v)) < 1e-14
    f.d2(dx, dv, do)
    assert rel(do.numpy(), oracle.bratu_d2(n, lam, x, v)) < 1e-14
    f.jacobi_diag(dx, do)
    assert rel(do.numpy(), oracle.bratu_jacobi_diag(n, lam, x)) < 1e-14


@pytest.mark.parametrize("n", SIZES)
def test_jvp_fd(ctx, oracle, n):
    import numa_hip as H
    x, v = _inputs(n)
    lam = 6.0
    f = H.Bratu3D(ctx, n, lam)
    F0 = oracle.bratu_eval(n, lam, x)
    dx, dF0, dv, do = ctx.array(x), ctx.array(F0), ctx.array(v), ctx.empty(n ** 3)
    f.jvp_fd(dx, dF0, dv, do)
    ref = oracle.bratu_jvp_fd(n, lam, x, v)
    assert rel(do.numpy(), ref) < 1e-7
    # mixed: v and out in FP32
    v32 = v.astype(np.float32)
    dv32, do32 = ctx.array(v32), ctx.empty(n ** 3, np.float32)
    f.jvp_fd(dx, dF0, dv32, do32)
    ref32 = oracle.bratu_jvp_fd(n, lam, x, v32.astype(np.float64))
    assert rel(do32.numpy(), ref32) < 2e-6


@pytest.mark.parametrize("n", SIZES)
def test_jvp_fd_rearranged(ctx, oracle, n):
    """The mixed solver's J v (v, E, out FP32) against the oracle's FP64
    forward difference (F(x+hv) - F(x))/h at the same FP32 v; E = lambda e^x
    to FP32 rounding."""
    x, v = _inputs(n)
    lam = 6.0
    f = H.Bratu3D(ctx, n, lam)
    v32 = v.astype(np.float32)
    dE, do = ctx.empty(n ** 3, np.float32), ctx.empty(n ** 3, np.float32)
    f.jvp_fd_rearranged(ctx.array(x), ctx.array(v32), dE, do)
    assert rel(dE.numpy(), lam * np.exp(x)) < 1e-7
    ref = oracle.bratu_jvp_fd(n, lam, x, v32.astype(np.float64))
    assert rel(do.numpy(), ref) < 2e-6


@pytest.mark.parametrize("hv", [5e-4, 9e-4, 2e-3, 3e-2])
def test_jvp_fd_rearranged_step_sizes(ctx, oracle, hv):
    """(e^{hv} - 1)/h on both sides of the Taylor form's bound |hv| <= 2^-10.
    x = -C except 0 at the spike of v, with C such that h v = hv there (h =
    sqrt(eps)(1 + ||x||)/||v||); lambda e^x vanishes off the spike, and
    lambda = 1e4 makes the exponential term most of J v at the spike, so its
    first-order factor 1 + hv/2 is far above the FP32 tolerance."""
    n = 24
    N = n ** 3
    k = 12 * n * n + 12 * n + 12
    v = 1e-3 * np.sin(np.arange(N, dtype=np.float64))
    v[k] = 1.0
    v32 = v.astype(np.float32)
    vnorm = np.linalg.norm(v32.astype(np.float64))
    xnorm = hv * vnorm / np.sqrt(np.finfo(float).eps) - 1.0
    x = np.full(N, -xnorm / np.sqrt(N - 1))
    x[k] = 0.0
    lam = 1e4
    f = H.Bratu3D(ctx, n, lam)
    dE, do = ctx.empty(N, np.float32), ctx.empty(N, np.float32)
    f.jvp_fd_rearranged(ctx.array(x), ctx.array(v32), dE, do)
    ref = oracle.bratu_jvp_fd(n, lam, x, v32.astype(np.float64))
    got = do.numpy()
    assert rel(got, ref) < 2e-6
    assert abs(got[k] - ref[k]) <= 2e-6 * abs(ref[k])
    # the exponential term is most of J v at the spike: a quotient without
    # its hv/2 term would miss by about hv/2 relative
    h = np.sqrt(np.finfo(float).eps) * (1.0 + np.linalg.norm(x)) / vnorm
    assert abs(h - hv) <= 1e-12 * hv
    expo = lam * np.expm1(h) / h
    assert expo > 0.5 * abs(ref[k])


def test_jvp_fd_zero_direction(ctx, oracle):
    import numa_hip as H
    n = 8
    x, _ = _inputs(n)
    f = H.Bratu3D(ctx, n, 1.0)
    F0 = oracle.bratu_eval(n, 1.0, x)
    do = ctx.array(np.full(n ** 3, 7.0))
    f.jvp_fd(ctx.array(x), ctx.array(F0), ctx.array(np.zeros(n ** 3)), do)
    assert np.all(do.numpy() == 0.0)  # operator.hpp:105-106


@pytest.mark.parametrize("n", [6, 8, 16, 31, 64])
def test_d2_fd(ctx, oracle, n):
    import numa_hip as H
    x, v = _inputs(n)
    v = 0.5 * v
    lam = 1.0
    f = H.Bratu3D(ctx, n, lam)
    F0 = oracle.bratu_eval(n, lam, x)
    do = ctx.empty(n ** 3)
    f.d2_fd(ctx.array(x), ctx.array(F0), ctx.array(v), do)
    ref = oracle.bratu_d2_fd(n, lam, x, v)
    assert rel(do.numpy(), ref) < 1e-4
    # FD truncation / cancellation against the analytic F''[v,v] is the same
    # on both sides; at the reference's own case n=6 it is below 1e-4
    # (test_operators.cpp:49-61).
    ana = oracle.bratu_d2(n, lam, x, v)
    assert rel(do.numpy(), ana) <= 1.5 * rel(ref, ana) + 1e-10
    if n == 6:
        assert rel(do.numpy(), ana) < 1e-4


@pytest.mark.parametrize("N", [1, 255, 4096, 100003, 2 ** 21 + 17])
def test_dot_axpy(ctx, oracle, N):
    rng = np.random.default_rng(N)
    a = rng.standard_normal(N)
    b = rng.standard_normal(N)
    import numa_hip as H
    da, db = ctx.array(a), ctx.array(b)
    r = C_dot(ctx, da, db, N)
    assert abs(r - np.dot(a, b)) <= 1e-12 * np.linalg.norm(a) * np.linalg.norm(b)
    H.check(ctx.lib.numa_hip_axpy(ctx.h, N, 0.75, da.ptr, db.ptr))
    assert np.array_equal(db.numpy(), b + 0.75 * a)
    # FP32 dot accumulates in FP64
    a32, b32 = a.astype(np.float32), b.astype(np.float32)
    r32 = C_dot(ctx, ctx.array(a32), ctx.array(b32), N, f32=True)
    assert abs(r32 - np.dot(a32.astype(np.float64), b32.astype(np.float64))) <= 1e-12 * N


def C_dot(ctx, da, db, N, f32=False):
    import ctypes as C
    import numa_hip as H
    out = C.c_double()
    f = ctx.lib.numa_hip_dot_f32 if f32 else ctx.lib.numa_hip_dot
    H.check(f(ctx.h, N, da.ptr, db.ptr, C.byref(out)))
    return out.value


def test_dot_deterministic(ctx):
    rng = np.random.default_rng(1)
    a = rng.standard_normal(3_000_001)
    da = ctx.array(a)
    vals = {C_dot(ctx, da, da, a.size) for _ in range(5)}
    assert len(vals) == 1


def _wave_tree(v):
    """A 64-lane wave sum as the kernels do it (lane l adds lane l + off for
    off = 32, 16, ..., 1; a lane past the wave adds its own value); lane 0."""
    v = v.copy()
    for off in (32, 16, 8, 4, 2, 1):
        sh = np.concatenate([v[..., off:], v[..., 64 - off:]], axis=-1)
        v = v + sh
    return v[..., 0]


def _finish(p):
    """The one-block finish of a reduction: 1024 strided sequential sums, then
    the wave sums of the 16 waves added in order."""
    K = 1024
    s = np.zeros(K)
    for t in range(K):
        acc = 0.0
        for a in p[t::K]:
            acc += a
        s[t] = acc
    ws = _wave_tree(s.reshape(16, 64))
    tot = 0.0
    for w in ws:
        tot += w
    return tot


@pytest.mark.parametrize("n", [33, 48, 64, 96, 100, 127])
def test_stencil_norm_sums_in_the_one_node_per_thread_order(ctx, n):
    """Below 128^3 the LDS-tiled stencil writes the partials of the
    one-node-per-thread form (64 x 4 tiles, one plane each, wave sums added
    over the 4 waves in order), so ||F|| is that form's total bit for bit:
    the reduction is redone on the host in exactly that order from the F the
    kernel wrote."""
    import numa_hip as H
    rng = np.random.default_rng(n)
    u = ctx.array(0.3 * rng.random(n ** 3))
    F = ctx.empty(n ** 3)
    nrm = H.Bratu3D(ctx, n, 6.0).eval(u, F)
    c = F.numpy() ** 2
    gx, gy4 = -(-n // 64), -(-n // 4)
    pad = np.zeros((n, gy4 * 4, gx * 64))
    pad[:, :n, :n] = c.reshape(n, n, n)
    waves = _wave_tree(pad.reshape(n, gy4, 4, gx, 64).transpose(0, 1, 3, 2, 4))
    part = np.zeros((n, gy4, gx))
    for w in range(4):
        part = part + waves[..., w]
    want = np.sqrt(_finish(part.reshape(-1)))
    assert nrm == want, (nrm, want)
    u.free()
    F.free()


def test_mfma_peak_is_plausible(ctx):
    """The measured dense MFMA throughput (a register-only loop of independent
    accumulator chains) lies below the published MI355X peaks and within a
    factor of three of them."""
    bf16 = ctx.mfma_peak(0)
    fp64 = ctx.mfma_peak(1)
    assert 2500.0 / 3 < bf16 <= 2500.0 * 1.05, bf16
    assert 78.6 / 3 < fp64 <= 78.6 * 1.05, fp64
