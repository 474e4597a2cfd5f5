"""The surrogate network restated in numpy (test infrastructure only).

Imported by tests/ as the checker of the HIP surrogate kernels; the product
path never imports it.  The network is n_in-256-256-1 with SiLU; the flat
FP32 weight vector is W1 (16 x 256) | b1 | W2 (256 x 256) | b2 | w3 | b3,
input index first.  Arithmetic is FP64; with ``rounding=True`` the values
the device holds in bf16 (inputs, W1, W2, first hidden layer, the two
backpropagated deltas that feed matrix products) are rounded to bf16 (round
to nearest even) at the same points, so the device's FP32-accumulated
results differ only by accumulation order.  ``rounding=False`` is the exact
network used to pin the backward pass against finite differences.
"""
from __future__ import annotations

import numpy as np

N_IN, HID = 16, 256
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_IN * HID
OFF_W2 = OFF_B1 + HID
OFF_B2 = OFF_W2 + HID * HID
OFF_W3 = OFF_B2 + HID
OFF_B3 = OFF_W3 + HID
N_WEIGHTS = OFF_B3 + 1


def bf16(x) -> np.ndarray:
    """Round to bf16 (nearest even) through FP32, returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    out = (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)
    return np.where(nan, np.nan, out)


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    return dict(W1=w[OFF_W1:OFF_B1].reshape(N_IN, HID), b1=w[OFF_B1:OFF_W2],
                W2=w[OFF_W2:OFF_B2].reshape(HID, HID), b2=w[OFF_B2:OFF_W3],
                w3=w[OFF_W3:OFF_B3], b3=w[OFF_B3])


def pack(g) -> np.ndarray:
    return np.concatenate([g["W1"].ravel(), g["b1"], g["W2"].ravel(), g["b2"],
                           g["w3"], [g["b3"]]])


def _sig(x):
    return 1.0 / (1.0 + np.exp(-x))


def silu(x):
    return x * _sig(x)


def silu_grad(x):
    s = _sig(x)
    return s * (1.0 + x * (1.0 - s))


def features(theta, mean, whiten, rounding=True):
    """Network input rows (n, 16): x = W (theta - mean), zero-padded."""
    theta = np.atleast_2d(np.asarray(theta, np.float64))
    n, d = theta.shape
    x = (theta - np.asarray(mean, np.float64)) @ np.asarray(whiten).reshape(d, d).T
    X = np.zeros((n, N_IN))
    X[:, :d] = x
    return bf16(X) if rounding else X


def forward(w, X, rounding=True):
    r = bf16 if rounding else (lambda v: np.asarray(v, np.float64))
    p = unpack(w)
    a1 = r(X) @ r(p["W1"]) + p["b1"]
    h1 = r(silu(a1))
    a2 = h1 @ r(p["W2"]) + p["b2"]
    h2 = silu(a2)
    s = h2 @ p["w3"] + p["b3"]
    return s, dict(X=r(X), a1=a1, h1=h1, a2=a2, h2=h2, s=s)


def loss_and_grad(w, X, t, rounding=True):
    """Mean squared error of s against standardised targets t and its
    gradient with respect to the flat weights."""
    r = bf16 if rounding else (lambda v: np.asarray(v, np.float64))
    p = unpack(w)
    s, c = forward(w, X, rounding)
    n = len(t)
    e = s - t
    loss = float(np.mean(e * e))
    g = 2.0 * e / n
    gw3 = c["h2"].T @ g
    gb3 = float(g.sum())
    d2 = g[:, None] * p["w3"][None, :] * silu_grad(c["a2"])
    gb2 = d2.sum(axis=0)
    d2r = r(d2)
    gW2 = c["h1"].T @ d2r
    d1 = (d2r @ r(p["W2"]).T) * silu_grad(c["a1"])
    gb1 = d1.sum(axis=0)
    gW1 = c["X"].T @ r(d1)
    return loss, pack(dict(W1=gW1, b1=gb1, W2=gW2, b2=gb2, w3=gw3, b3=gb3))
