"""CPU (numpy, FP64) oracle of the learned preconditioner's 3D conv net.

TEST INFRASTRUCTURE ONLY: used by tests/ to check the HIP forward and
backward kernels; never a product path.  The network has no counterpart in
the reference (kapumota/chebyshev-nonlinear-lab has no learned components);
this file is the specification the kernels are checked against, and it is
itself pinned by tests/test_oracle_convnet.py (scipy.ndimage.correlate for
the convolution, central differences for every gradient, closed forms of
GELU, torch's bfloat16 conversion for the rounding).

Network: L layers of 3x3x3 convolutions with zero padding (the zero
Dirichlet boundary) on an n^3 grid, GELU between layers, channels
[1, C, ..., C, 1].  Activations are channel-major a[c, i, j, k]; weights W[l]
have shape (Cout, Cin, 3, 3, 3) with W[co, ci, di+1, dj+1, dk+1] multiplying
a[ci, i+di, j+dj, k+dk].

    pre_l = conv(a_{l-1}, W_l) + b_l,   a_l = gelu(pre_l) for l < L, out = pre_L

GELU is the exact form 0.5 x (1 + erf(x / sqrt 2)).  Loss of a batch:
    "correction": L = mean_b ||out_b - y_b||^2 / ||y_b||^2
    "residual":   L = mean_b ||A7 (out_b - y_b)||^2 / ||A7 y_b||^2
with A7 u = 6u - (sum of the six neighbours), zero outside.

bf16=True reproduces where the device rounds to bfloat16 (round to nearest
even): the weights of the inner (C -> C) layers, every stored activation
a_l = bf16(gelu(pre_l)), the stored pre-activations used by gelu', and the
back-propagated gradients dL/dpre_l of the inner layers.  The first and last
layers keep FP32 weights on the device (FP64 here).
"""
from __future__ import annotations

import math

import numpy as np

OFFSETS = [(di, dj, dk) for di in (-1, 0, 1) for dj in (-1, 0, 1) for dk in (-1, 0, 1)]


try:  # vectorised erf (cephes); math.erf elementwise otherwise
    from scipy.special import erf as _erf
except ImportError:  # pragma: no cover
    def _erf(x):
        return np.vectorize(math.erf, otypes=[np.float64])(x)


def gelu(x):
    return 0.5 * x * (1.0 + _erf(x / math.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + _erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2 * math.pi)


def bf16(a):
    """Round to bfloat16 (nearest, ties to even) through float32; returns float64."""
    f = np.ascontiguousarray(np.asarray(a, dtype=np.float64).astype(np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64).reshape(np.shape(a))


def shift(a, d):
    """s[..., i, j, k] = a[..., i+di, j+dj, k+dk] with zeros outside."""
    n = a.shape[-1]
    out = np.zeros_like(a)
    src, dst = [], []
    for dd in d:
        if dd >= 0:
            dst.append(slice(0, n - dd))
            src.append(slice(dd, n))
        else:
            dst.append(slice(-dd, n))
            src.append(slice(0, n + dd))
    out[(Ellipsis,) + tuple(dst)] = a[(Ellipsis,) + tuple(src)]
    return out


def conv(a, W, b):
    """a: (Cin, n, n, n); W: (Cout, Cin, 3, 3, 3); b: (Cout,)"""
    Cout = W.shape[0]
    n = a.shape[-1]
    out = np.broadcast_to(np.asarray(b, np.float64).reshape(Cout, 1, 1, 1), (Cout, n, n, n)).copy()
    for d in OFFSETS:
        w = W[:, :, d[0] + 1, d[1] + 1, d[2] + 1]
        out += np.einsum("oc,cijk->oijk", w, shift(a, d))
    return out


def conv_backward(a, W, dout):
    """Gradients of conv(a, W, b) w.r.t. a, W and b given dL/dout."""
    dW = np.zeros_like(W)
    da = np.zeros_like(a)
    for d in OFFSETS:
        dW[:, :, d[0] + 1, d[1] + 1, d[2] + 1] = np.einsum("oijk,cijk->oc", dout, shift(a, d))
        w = W[:, :, d[0] + 1, d[1] + 1, d[2] + 1]
        # out[p] += w a[p + d]  =>  da[q] += w^T dout[q - d]
        da += np.einsum("oc,oijk->cijk", w, shift(dout, (-d[0], -d[1], -d[2])))
    db = dout.sum(axis=(1, 2, 3))
    return da, dW, db


NB6 = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]


def a7(u):
    """6 u - (sum of the six neighbours), zero outside (symmetric)."""
    return 6.0 * u - sum(shift(u, d) for d in NB6)


def _inner(l, L):
    return 0 < l < L - 1


def forward(x, Ws, bs, bf16_mode=False):
    """x: (n, n, n) input; returns (out (n,n,n), (pres, acts))."""
    L = len(Ws)
    a = np.asarray(x, np.float64)[None]
    pres, acts = [], [a]
    for l, (W, b) in enumerate(zip(Ws, bs)):
        Wl = bf16(W) if (bf16_mode and _inner(l, L)) else W
        pre = conv(a, Wl, b)
        if l + 1 < L:
            a = gelu(pre)
            if bf16_mode:
                a = bf16(a)
                pre = bf16(pre)
        else:
            a = pre
        pres.append(pre)
        acts.append(a)
    return a[0], (pres, acts)


def loss_and_grads(xs, ys, Ws, bs, bf16_mode=False, loss_kind="correction"):
    """Batch loss and gradients.  xs, ys: (B, n, n, n)."""
    L = len(Ws)
    B = len(xs)
    loss = 0.0
    dWs = [np.zeros_like(W) for W in Ws]
    dbs = [np.zeros(np.shape(b)) for b in bs]
    outs = []
    for x, y in zip(xs, ys):
        out, (pres, acts) = forward(x, Ws, bs, bf16_mode)
        outs.append(out)
        if loss_kind == "correction":
            tn = float(np.sum(y * y))
            loss += float(np.sum((out - y) ** 2)) / tn / B
            g = (2.0 / (B * tn)) * (out - y)[None]
        else:
            ay = a7(y)
            tn = float(np.sum(ay * ay))
            e = a7(out - y)
            loss += float(np.sum(e * e)) / tn / B
            g = (2.0 / (B * tn)) * a7(e)[None]
        for l in reversed(range(L)):
            if l + 1 < L:
                g = g * gelu_grad(pres[l])
                if bf16_mode:
                    g = bf16(g)
            Wl = bf16(Ws[l]) if (bf16_mode and _inner(l, L)) else Ws[l]
            da, dW, db = conv_backward(acts[l], Wl, g)
            dWs[l] += dW
            dbs[l] += db
            g = da
    return loss, dWs, dbs, np.array(outs)


# ---- flat parameter layout of the device network ---------------------------

def layer_shapes(C=24, L=6):
    ch = [1] + [C] * (L - 1) + [1]
    return [(ch[l + 1], ch[l], 3, 3, 3) for l in range(L)]


def unflatten(flat, C=24, L=6):
    """Flat device parameters -> (Ws, bs): per layer W (row-major) then b."""
    Ws, bs, off = [], [], 0
    for shp in layer_shapes(C, L):
        cnt = int(np.prod(shp))
        Ws.append(np.asarray(flat[off:off + cnt], np.float64).reshape(shp))
        off += cnt
        bs.append(np.asarray(flat[off:off + shp[0]], np.float64))
        off += shp[0]
    assert off == len(flat)
    return Ws, bs


def flatten(Ws, bs):
    return np.concatenate([np.concatenate([W.ravel(), np.ravel(b)]) for W, b in zip(Ws, bs)])
