"""CPU oracle of the property surrogate (TEST INFRASTRUCTURE ONLY).

A numpy restatement of the network that the device library trains and runs
for THERM_PROP_UPDATE: 1 -> 64 -> 64 -> 5, tanh, weights and activations as
pairs of BF16 values, BF16 products with FP32 accumulation.  Only tests/ load it.  The reference has no surrogate; what it
computes is fixed by the property table it replaces (the labels), so the
network's definition below is this build's own and is pinned by

* bf16 rounding (round to nearest even) against torch's bfloat16 cast;
* the backward pass against central finite differences of the FP64 loss
  (bf16 rounding switched off, so the loss is differentiable);

and the device kernels are checked against this oracle.

Parameters (flat float32 vector, NPAR = 4613):
    w1[64] b1[64] W2[64][64] b2[64] W3[5][64] b3[5]      (W[out][in])
Forward of one sample h (normalisation hc, hs, m[5], s[5]), with bf()
rounding to bfloat16, hi(v) = bf(v), lo(v) = bf(v - hi(v)), q(w) = hi(w) + lo(w):
    x  = (h - hc) * hs
    a1 = tanh(q(w1) * x + b1)
    a2 = tanh(hi(W2) @ (hi(a1) + lo(a1)) + lo(W2) @ hi(a1) + b2)
    o  = hi(W3) @ (hi(a2) + lo(a2)) + lo(W3) @ hi(a2) + b3
    y  = m + s * o            (T, rho, mu, kappa, c_p)
where W stands for q(W) (the stored weight).
Loss over a batch of B samples with table values t:
    L = 1/(5B) sum_s sum_k ((y_k - t_k) / t_k)^2
Gradients are taken with respect to the stored weights q(W) (the backward
uses q(W) as the weight) and applied to the FP32 master weights
(straight-through).
"""
from __future__ import annotations

import numpy as np

H1, H2, NOUT = 64, 64, 5
O_W1, O_B1, O_W2 = 0, 64, 128
O_B2 = O_W2 + H2 * H1
O_W3 = O_B2 + H2
O_B3 = O_W3 + NOUT * H2
NPAR = O_B3 + NOUT


def bf16(a):
    """Round float32 values to bfloat16 (nearest, ties to even); returns float32."""
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    nan = np.isnan(a)
    if nan.any():
        out = np.where(nan, a, out)
    return out


def split(p):
    return (p[O_W1:O_B1], p[O_B1:O_W2], p[O_W2:O_B2].reshape(H2, H1), p[O_B2:O_W3],
            p[O_W3:O_B3].reshape(NOUT, H2), p[O_B3:NPAR])


def hilo(v):
    """(hi, lo) bfloat16 pair of float32 values, as float64."""
    v = np.asarray(v, np.float32)
    h = bf16(v)
    return h.astype(np.float64), bf16(v - h).astype(np.float64)


def qw(w):
    """The stored weight q(w) = hi(w) + lo(w) (exact in float32)."""
    h, l = hilo(w)
    return h + l


def forward(p, norm, h, rounding=True):
    """Network outputs y[B, 5] and the activations (float64 arithmetic on the
    float32 operands)."""
    hc, hs, m, s = norm
    w1, b1, W2, b2, W3, b3 = split(np.asarray(p, np.float32))
    b1, b2, b3 = (np.asarray(v, np.float64) for v in (b1, b2, b3))
    x = ((np.asarray(h, np.float64) - hc) * hs).astype(np.float32).astype(np.float64)

    def mm(a, W):
        if not rounding:
            return a @ np.asarray(W, np.float64).T
        Wh, Wl = hilo(qw(W))
        ah, al = hilo(a)
        return (ah + al) @ Wh.T + ah @ Wl.T

    w1q = qw(w1) if rounding else np.asarray(w1, np.float64)
    a1 = np.tanh(np.outer(x, w1q) + b1)                                  # [B, 64]
    if rounding:
        a1 = a1.astype(np.float32).astype(np.float64)
    a2 = np.tanh(mm(a1, W2) + b2)
    if rounding:
        a2 = a2.astype(np.float32).astype(np.float64)
    o = mm(a2, W3) + b3
    y = np.asarray(m, np.float64) + np.asarray(s, np.float64) * o
    return y, (x, a1, a2, o)


def loss_grad(p, norm, h, t, rounding=True):
    """Loss and its gradient (float64) with respect to the (rounded) parameters."""
    hc, hs, m, s = norm
    p = np.asarray(p, np.float32)
    w1, b1, W2, b2, W3, b3 = split(p)
    rq = qw if rounding else (lambda v: np.asarray(v, np.float64))
    y, (x, a1, a2, o) = forward(p, norm, h, rounding)
    t = np.asarray(t, np.float64)
    B = len(x)
    e = (y - t) / t
    L = float((e * e).sum() / (5 * B))
    dy = 2.0 * e / t / (5 * B)
    do = dy * np.asarray(s, np.float64)                                  # [B, 5]
    g = np.zeros(NPAR)
    g[O_B3:NPAR] = do.sum(0)
    g[O_W3:O_B3] = (do.T @ a2).ravel()
    dz2 = (do @ rq(W3)) * (1.0 - a2 * a2)
    g[O_B2:O_W3] = dz2.sum(0)
    g[O_W2:O_B2] = (dz2.T @ a1).ravel()
    dz1 = (dz2 @ rq(W2)) * (1.0 - a1 * a1)
    g[O_B1:O_W2] = dz1.sum(0)
    g[O_W1:O_B1] = (dz1 * x[:, None]).sum(0)
    return L, g
