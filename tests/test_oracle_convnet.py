"""Pins of the conv-net oracle (oracle/convnet_ref.py) on the CPU.

The learned preconditioner has no counterpart in the reference, so the oracle
is pinned against independent routines: scipy.ndimage.correlate for the
convolution, the adjoint (dot-product) identity for the transposed
convolution, central differences for every parameter gradient, the normal
CDF for GELU and torch's float32 -> bfloat16 conversion for the rounding.
"""
import ctypes as C
import math

import numpy as np
import pytest
import scipy.ndimage as ndi
import scipy.stats as sps

from conftest import ORACLE_DIR  # noqa: F401  (puts oracle/ on sys.path)
import convnet_ref as cr


def _rand_net(rng, C_=3, L=4, scale=0.3):
    Ws, bs = [], []
    for shp in cr.layer_shapes(C_, L):
        Ws.append(rng.normal(0, scale, shp))
        bs.append(rng.normal(0, 0.1, shp[0]))
    return Ws, bs


def test_conv_matches_scipy_correlate():
    rng = np.random.default_rng(1)
    a = rng.normal(size=(3, 7, 7, 7))
    W = rng.normal(size=(4, 3, 3, 3, 3))
    b = rng.normal(size=4)
    out = cr.conv(a, W, b)
    for co in range(4):
        ref = b[co] + sum(ndi.correlate(a[ci], W[co, ci], mode="constant", cval=0.0)
                          for ci in range(3))
        np.testing.assert_allclose(out[co], ref, rtol=1e-12, atol=1e-12)


def test_conv_backward_adjoint_identity():
    rng = np.random.default_rng(2)
    a = rng.normal(size=(3, 6, 6, 6))
    W = rng.normal(size=(2, 3, 3, 3, 3))
    g = rng.normal(size=(2, 6, 6, 6))
    da, dW, db = cr.conv_backward(a, W, g)
    # <conv(a, W, 0), g> = <a, da> = <W, dW>
    lhs = float(np.sum(cr.conv(a, W, np.zeros(2)) * g))
    assert abs(lhs - float(np.sum(a * da))) <= 1e-10 * abs(lhs)
    assert abs(lhs - float(np.sum(W * dW))) <= 1e-10 * abs(lhs)
    np.testing.assert_allclose(db, g.sum(axis=(1, 2, 3)), rtol=1e-14)


def test_gelu_closed_forms():
    x = np.linspace(-6, 6, 97)
    np.testing.assert_allclose(cr.gelu(x), x * sps.norm.cdf(x), rtol=1e-13, atol=1e-15)
    assert cr.gelu(np.array([0.0]))[0] == 0.0
    assert abs(cr.gelu_grad(np.array([0.0]))[0] - 0.5) < 1e-15
    h = 1e-6
    fd = (cr.gelu(x + h) - cr.gelu(x - h)) / (2 * h)
    np.testing.assert_allclose(cr.gelu_grad(x), fd, rtol=1e-8, atol=1e-9)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.normal(size=2000) * 10.0 ** rng.integers(-8, 8, 2000),
                        [0.0, 1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, -2.5, 65504.0]])
    ref = torch.tensor(x.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(cr.bf16(x), ref)


def test_a7_is_the_symmetric_7_point_operator():
    rng = np.random.default_rng(8)
    u = rng.normal(size=(5, 5, 5))
    v = rng.normal(size=(5, 5, 5))
    assert abs(np.sum(cr.a7(u) * v) - np.sum(u * cr.a7(v))) < 1e-12
    ref = ndi.correlate(u, np.array([[[0, 0, 0], [0, -1, 0], [0, 0, 0]],
                                     [[0, -1, 0], [-1, 6, -1], [0, -1, 0]],
                                     [[0, 0, 0], [0, -1, 0], [0, 0, 0]]], float),
                        mode="constant", cval=0.0)
    np.testing.assert_allclose(cr.a7(u), ref, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("loss_kind", ["correction", "residual"])
def test_gradients_central_differences(loss_kind):
    rng = np.random.default_rng(4)
    L, C_, n, B = 4, 3, 5, 2
    Ws, bs = _rand_net(rng, C_, L)
    xs = rng.normal(size=(B, n, n, n))
    ys = rng.normal(size=(B, n, n, n))
    lg = lambda W, b: cr.loss_and_grads(xs, ys, W, b, False, loss_kind)  # noqa: E731
    loss, dWs, dbs, _ = lg(Ws, bs)
    h = 1e-6
    for l in range(L):
        for _ in range(6):
            idx = tuple(int(rng.integers(0, s)) for s in Ws[l].shape)
            Wp = [W.copy() for W in Ws]
            Wm = [W.copy() for W in Ws]
            Wp[l][idx] += h
            Wm[l][idx] -= h
            fd = (lg(Wp, bs)[0] - lg(Wm, bs)[0]) / (2 * h)
            assert abs(fd - dWs[l][idx]) <= 1e-6 * max(1e-3, abs(fd)), (l, idx, fd, dWs[l][idx])
        c = int(rng.integers(0, bs[l].size))
        bp = [b.copy() for b in bs]
        bm = [b.copy() for b in bs]
        bp[l][c] += h
        bm[l][c] -= h
        fd = (lg(Ws, bp)[0] - lg(Ws, bm)[0]) / (2 * h)
        assert abs(fd - dbs[l][c]) <= 1e-6 * max(1e-3, abs(fd)), (l, c, fd, dbs[l][c])


def test_residual_loss_closed_form():
    rng = np.random.default_rng(6)
    Ws, bs = _rand_net(rng)
    xs = rng.normal(size=(2, 4, 4, 4))
    ys = rng.normal(size=(2, 4, 4, 4))
    loss, _, _, outs = cr.loss_and_grads(xs, ys, Ws, bs, loss_kind="residual")
    ref = np.mean([np.sum(cr.a7(o - y) ** 2) / np.sum(cr.a7(y) ** 2) for o, y in zip(outs, ys)])
    assert abs(loss - ref) <= 1e-14 * ref


def test_loss_is_relative_mse():
    rng = np.random.default_rng(5)
    Ws, bs = _rand_net(rng)
    xs = rng.normal(size=(2, 4, 4, 4))
    ys = rng.normal(size=(2, 4, 4, 4))
    loss, _, _, outs = cr.loss_and_grads(xs, ys, Ws, bs)
    ref = np.mean([np.sum((o - y) ** 2) / np.sum(y * y) for o, y in zip(outs, ys)])
    assert abs(loss - ref) <= 1e-14 * ref
    # a network whose output is the target has zero loss: last layer = identity tap
    Wz = [W * 0 for W in Ws]
    bz = [b * 0 for b in bs]
    Wz[0][0, 0, 1, 1, 1] = 1.0          # a_1[0] = gelu(x)
    for l in range(1, 3):
        Wz[l][0, 0, 1, 1, 1] = 1.0      # gelu(gelu(...)) chain
    Wz[3][0, 0, 1, 1, 1] = 1.0
    out, _ = cr.forward(xs[0], Wz, bz)
    y = xs[0]
    for _ in range(3):
        y = cr.gelu(y)
    np.testing.assert_allclose(out, y, rtol=1e-14, atol=1e-15)


def test_flat_layout_matches_library():
    """Parameter count of the flat layout agrees with the C ABI (no GPU call)."""
    import numa_hip as nh
    lib = nh.load()
    cnt = C.c_int64()
    nh.check(lib.numa_hip_net_param_count(C.byref(cnt)))
    shapes = cr.layer_shapes()
    total = sum(int(np.prod(s)) + s[0] for s in shapes)
    assert cnt.value == total == 63625
    flat = np.arange(total, dtype=np.float64)
    Ws, bs = cr.unflatten(flat)
    np.testing.assert_array_equal(cr.flatten(Ws, bs), flat)
    assert Ws[1][2, 3, 0, 1, 2] == 648 + 24 + (2 * 24 + 3) * 27 + 0 * 9 + 1 * 3 + 2


def test_device_gelu_formula():
    """The device GELU evaluates Phi(x) by Abramowitz & Stegun 7.1.26
    (erfc(z) = (a1 t + ... + a5 t^5) exp(-z^2), t = 1/(1 + p z)); restated here
    in FP64 it stays within 1.5e-7 of the exact Phi (the published bound) over
    [-10, 10], and so does its GELU' = Phi + x phi."""
    import math
    p, a = 0.3275911, (0.254829592, -0.284496736, 1.421413741, -1.453152027, 1.061405429)
    x = np.linspace(-10.0, 10.0, 400001)
    z = np.abs(x) / math.sqrt(2.0)
    t = 1.0 / (1.0 + p * z)
    poly = t * (a[0] + t * (a[1] + t * (a[2] + t * (a[3] + t * a[4]))))
    e = np.exp(-z * z)
    phi = np.where(x >= 0.0, 1.0 - 0.5 * poly * e, 0.5 * poly * e)
    exact = 0.5 * (1.0 + np.vectorize(math.erf)(x / math.sqrt(2.0)))
    assert np.max(np.abs(phi - exact)) < 1.5e-7
    grad = phi + x * e / math.sqrt(2.0 * math.pi)
    assert np.max(np.abs(grad - cr.gelu_grad(x))) < 1.5e-7
    assert np.max(np.abs(x * phi - cr.gelu(x))) < 1.5e-6


@pytest.mark.parametrize("name", ["convnet_200k_n16-32", "convnet_200k_n64-128"])
def test_committed_weights_load(name):
    """The committed parameter sets are plain FP32 arrays in the flat layout
    (63,625 values, read without unpickling), their training record names the
    200k-pair budget, and the oracle's forward pass over them is finite."""
    import json
    import os
    import numa_hip as nh
    path = os.path.join(os.path.dirname(nh.__file__), "weights", name)
    w = np.load(path + ".npy")
    assert w.dtype == np.float32 and w.shape == (63625,)
    assert np.all(np.isfinite(w)) and np.abs(w).max() < 100.0
    rec = json.load(open(path + ".json"))
    assert rec["pairs"] == 200000 and rec["loss"] == "residual"
    assert rec["loss_last_window"] < rec["loss_first_window"]
    Ws, bs = cr.unflatten(w.astype(np.float64))
    x = np.random.default_rng(3).normal(size=(8, 8, 8))
    y = cr.forward(x, Ws, bs)[0]
    assert y.shape == x.shape and np.all(np.isfinite(y)) and np.abs(y).max() > 0.0
