"""TEST INFRASTRUCTURE ONLY: float64 numpy reference of the field surrogate
network (hpsim_amd/csrc/hps_surrogate.hip), for checking the GPU forward
pass and gradients.

Layers come from the network's layer table (rows {Cin, Cout, taps, dilation,
relu, residual, weight offset, bias offset}); weights are [Cout][taps][Cin]
(taps row-major over (dy, dx) in {-1, 0, 1}^2 times the dilation), biases
[Cout]; activations NHWC.  The loss is the per-sample relative L2 error of
output channels 0, 1; the gradient is that of the mean over samples.
"""
from __future__ import annotations

import numpy as np


def _taps(taps, dil):
    if taps == 1:
        return [(0, 0)]
    return [((t // 3 - 1) * dil, (t % 3 - 1) * dil) for t in range(9)]


def _shift(x, dy, dx):
    """y[b, i, j] = x[b, i + dy, j + dx], zero outside."""
    B, H, W, C = x.shape
    y = np.zeros_like(x)
    i0, i1 = max(0, -dy), min(H, H - dy)
    j0, j1 = max(0, -dx), min(W, W - dx)
    if i0 < i1 and j0 < j1:
        y[:, i0:i1, j0:j1] = x[:, i0 + dy:i1 + dy, j0 + dx:j1 + dx]
    return y


def _conv(x, w, b, taps, dil):
    out = np.zeros(x.shape[:3] + (w.shape[0],))
    for t, (dy, dx) in enumerate(_taps(taps, dil)):
        out += _shift(x, dy, dx) @ w[:, t, :].T
    return out + b


def forward(params, table, x):
    """Returns (prediction [B][H][W][Cout_last], cache)."""
    x = np.asarray(x, np.float64)
    cache = []
    for (cin, cout, taps, dil, relu, resid, wo, bo) in table:
        w = params[wo:wo + cout * taps * cin].reshape(cout, taps, cin).astype(np.float64)
        b = params[bo:bo + cout].astype(np.float64)
        z = _conv(x, w, b, taps, dil)
        a = np.maximum(z, 0.0) if relu else z
        y = x + a if resid else a
        cache.append((x, a, w))
        x = y
    return x, cache


def loss_and_grad(params, table, x, tgt):
    """Per-sample losses and the gradient of their mean."""
    pred, cache = forward(params, table, x)
    tgt = np.asarray(tgt, np.float64)
    B = x.shape[0]
    e = pred[..., :2] - tgt
    den = (tgt ** 2).reshape(B, -1).sum(1) + 1e-30
    loss = (e ** 2).reshape(B, -1).sum(1) / den
    d = np.zeros_like(pred)
    d[..., :2] = 2.0 * e / den[:, None, None, None] / B
    grad = np.zeros(params.size)
    for k in range(len(table) - 1, -1, -1):
        cin, cout, taps, dil, relu, resid, wo, bo = table[k]
        xin, a, w = cache[k]
        dz = d * (a > 0) if relu else d
        gw = np.zeros((cout, taps, cin))
        dx_ = np.zeros_like(xin)
        for t, (dy, dxs) in enumerate(_taps(taps, dil)):
            xs = _shift(xin, dy, dxs)
            gw[:, t, :] = np.einsum("bhwo,bhwi->oi", dz, xs)
            dx_ += _shift(dz @ w[:, t, :], -dy, -dxs)
        grad[wo:wo + cout * taps * cin] = gw.ravel()
        grad[bo:bo + cout] = dz.reshape(-1, cout).sum(0)
        d = dx_ + (d if resid else 0.0)
    return loss, grad
