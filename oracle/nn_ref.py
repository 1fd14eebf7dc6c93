"""TEST INFRASTRUCTURE ONLY: float64 numpy reference of the SCHEFF field
surrogate network (hpsim_amd/csrc/hps_unet.hip), for checking the GPU
forward pass and gradients.  Nothing in the product imports it.

The network is a one-dimensional convolutional U-Net along the bunch axis z,
with the source rings as channels: input x[b][z][rs] (z < Z0), output
y[b][z][c].  Each layer row of the table is
    (mode, Cin, Cout, taps, dilation, skip, weight offset, film offset)
with weights W[Cout][taps][Cin] and FiLM coefficients A[Cout][3] in the flat
parameter vector.  A layer computes
    pre[b][zo][co] = sum_{t, ci} W[co][t][ci] x[b][zi(zo, t)][ci]
    y = pre * (1 + sum_j A[co][j] f_j(b)) [+ output of layer `skip`]
with f(b) = (v, v^2, v^3), v = 2 tanh(u_b / 2), the conditioning features of
sample b (u_b the standardised conditioning, bounded by the tanh) and the
source index zi of the mode (terms with zi outside [0, Zin) vanish):
    0 same   (3 taps)  zi = zo + (t - 1) * dilation
    1 point  (1 tap)   zi = zo
    2 down   (2 taps)  zi = 2 zo + t                 (Zout = Zin / 2)
    3 up     (2 taps)  zi = zo // 2 when zo % 2 == t (Zout = 2 Zin)
Every layer reads the previous layer's output (layer 0 the input).  There is
no bias and no activation: the field is linear in the charge, and the
conditioning enters only through the FiLM scales.  The loss is the mean over
samples of |y - target|^2 / |target|^2 over the first `cout` channels of the
first `zn` positions.
"""
from __future__ import annotations

import numpy as np

SAME, POINT, DOWN, UP = 0, 1, 2, 3
TAPS = {SAME: 3, POINT: 1, DOWN: 2, UP: 2}


def zout_of(mode, zin):
    return zin // 2 if mode == DOWN else (2 * zin if mode == UP else zin)


def _src(mode, zo, t, dil, zin):
    """Source positions for output positions zo (array) and tap t; -1: none."""
    if mode == SAME:
        zi = zo + (t - 1) * dil
    elif mode == POINT:
        zi = zo.copy()
    elif mode == DOWN:
        zi = 2 * zo + t
    else:
        zi = np.where(zo % 2 == t, zo // 2, -1)
    return np.where((zi >= 0) & (zi < zin), zi, -1)


def conv(x, w, mode, dil):
    """pre[b][zo][co] of one layer; x [B][Zin][Cin], w [Cout][taps][Cin]."""
    B, zin, _ = x.shape
    zo = np.arange(zout_of(mode, zin))
    out = np.zeros((B, zo.size, w.shape[0]))
    for t in range(w.shape[1]):
        zi = _src(mode, zo, t, dil, zin)
        ok = zi >= 0
        out[:, ok, :] += x[:, zi[ok], :] @ w[:, t, :].T
    return out


def conv_grads(x, w, mode, dil, dpre):
    """(dx, dw) of conv for the output gradient dpre."""
    B, zin, _ = x.shape
    zo = np.arange(dpre.shape[1])
    dx = np.zeros_like(x)
    dw = np.zeros_like(w)
    for t in range(w.shape[1]):
        zi = _src(mode, zo, t, dil, zin)
        ok = zi >= 0
        dw[:, t, :] += np.einsum("bzo,bzi->oi", dpre[:, ok, :], x[:, zi[ok], :])
        np.add.at(dx, (slice(None), zi[ok], slice(None)), dpre[:, ok, :] @ w[:, t, :])
    return dx, dw


def _unpack(params, row):
    mode, cin, cout, taps, dil, skip, w_off, a_off = row
    w = params[w_off:w_off + cout * taps * cin].reshape(cout, taps, cin)
    a = params[a_off:a_off + cout * 3].reshape(cout, 3)
    return w, a


def features(u):
    v = 2.0 * np.tanh(np.asarray(u, dtype=np.float64) / 2.0)
    return np.stack([v, v * v, v * v * v], axis=1)


def forward(params, table, x, u, keep=False):
    """Output of the last layer; with keep, also the per-layer (input, pre,
    scale, output) for the backward pass."""
    f = features(u)
    h = x
    outs, saved = [], []
    for row in table:
        mode, _, _, _, dil, skip = row[:6]
        w, a = _unpack(params, row)
        pre = conv(h, w, mode, dil)
        s = 1.0 + f @ a.T                                  # [B][Cout]
        y = pre * s[:, None, :]
        if skip >= 0:
            y = y + outs[skip]
        saved.append((h, pre, s))
        outs.append(y)
        h = y
    return (h, saved) if keep else h


def loss_and_grad(params, table, x, u, tgt, zn, cout):
    """Per-sample losses and the gradient of their mean."""
    f = features(u)
    y, saved = forward(params, table, x, u, keep=True)
    B = x.shape[0]
    d = y[:, :zn, :cout] - tgt
    nt = (tgt * tgt).sum(axis=(1, 2))
    loss = (d * d).sum(axis=(1, 2)) / nt
    grad = np.zeros_like(params)
    dys = [None] * len(table)
    dy = np.zeros_like(y)
    dy[:, :zn, :cout] = 2.0 * d / nt[:, None, None] / B
    dys[-1] = dy
    for k in range(len(table) - 1, -1, -1):
        row = table[k]
        mode, _, cout_k, taps, dil, skip, w_off, a_off = row
        w, a = _unpack(params, row)
        h, pre, s = saved[k]
        dy = dys[k]
        if skip >= 0:
            dys[skip] = dy.copy() if dys[skip] is None else dys[skip] + dy
        g = (dy * pre).sum(axis=1)                          # [B][Cout]
        grad[a_off:a_off + cout_k * 3] += (g.T @ f).ravel()
        dpre = dy * s[:, None, :]
        dx, dw = conv_grads(h, w, mode, dil, dpre)
        grad[w_off:w_off + w.size] += dw.ravel()
        if k > 0:
            dys[k - 1] = dx if dys[k - 1] is None else dys[k - 1] + dx
    return loss, grad


def bf16(a):
    """Round to bfloat16 (nearest even) and back to float64."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)
