"""numpy checker of the learned-lifting kernels (test infrastructure).

* mlp / mlp_grad: the network on bf16-rounded weights, inputs and
  activations with fp64 arithmetic, and its backpropagation (straight through
  the roundings), to check the GPU's MFMA evaluation and gradient kernel.
* integer_lifting: the integer interpolet (+ CDF 5/3 update) transform along
  every axis, restated with numpy integer arithmetic, to check the learned
  transform's fallback path bit for bit.
"""
from __future__ import annotations

import numpy as np

H, I = 32, 8
W1, B1 = 0, 256
W2 = B1 + 32
B2 = W2 + 1024
W3 = B2 + 32
B3 = W3 + 32
P = B3 + 1


def bf16(x):
    x = np.asarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (r.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


# The kernels' GELU: z * g(t), t = clip(z, -c, c), g = clip(0.5 + t P(t^2), 0, 1)
# (paliwa_amd/csrc/learned.hip), a fit to the tanh form; test_learned_ref.py
# pins it against that form and its derivative against differences.
GELU_CLAMP = np.float32(3.4513)
GELU_P = np.array([3.93313942e-01, -5.90118521e-02, 6.50023659e-03, -3.92342173e-04,
                   9.70034432e-06], dtype=np.float32).astype(np.float64)
GELU_CLAMP64 = float(GELU_CLAMP)


def gelu_tanh(z):
    """The tanh form the kernels' GELU approximates."""
    return 0.5 * z * (1.0 + np.tanh(0.7978845608028654 * (z + 0.044715 * z ** 3)))


def _gelu_parts(z):
    z = np.asarray(z, dtype=np.float64)
    t = np.clip(z, -GELU_CLAMP64, GELU_CLAMP64)
    u = t * t
    p = np.polyval(GELU_P[::-1], u)
    raw = 0.5 + t * p
    return z, t, u, p, raw


def gelu(z):
    z, t, u, p, raw = _gelu_parts(z)
    return z * np.clip(raw, 0.0, 1.0)


def gelu_grad(z):
    z, t, u, p, raw = _gelu_parts(z)
    g = np.clip(raw, 0.0, 1.0)
    dp = np.polyval((GELU_P[1:] * np.arange(1, 5))[::-1], u)
    dg = p + 2.0 * u * dp
    inside = (np.abs(z) < GELU_CLAMP64) & (raw > 0.0) & (raw < 1.0)
    return np.where(inside, g + z * dg, g)


def unpack(w):
    w = np.asarray(w, dtype=np.float32)
    return (w[W1:B1].reshape(H, I), w[B1:W2], w[W2:B2].reshape(H, H), w[B2:W3],
            w[W3:B3], w[B3])


def mlp(u, w, rounding=True):
    """Outputs on inputs u [n, 8] (rounded as the GPU rounds them)."""
    rnd = bf16 if rounding else (lambda v: np.asarray(v, dtype=np.float64))
    w1, b1, w2, b2, w3, b3 = unpack(w)
    x = rnd(u)
    z1 = x @ rnd(w1).T + rnd(b1)
    h1 = rnd(gelu(z1))
    z2 = h1 @ rnd(w2).T + b2.astype(np.float64)
    h2 = rnd(gelu(z2))
    return h2 @ rnd(w3) + float(b3)


def mlp_grad(u, t, w, rounding=True):
    """(loss, gradient) of mean (mlp(u) - t)^2, fp64; with rounding, straight
    through the bf16 roundings."""
    rnd = bf16 if rounding else (lambda v: np.asarray(v, dtype=np.float64))
    w1, b1, w2, b2, w3, b3 = unpack(w)
    n = u.shape[0]
    x = rnd(u)
    z1 = x @ rnd(w1).T + rnd(b1)
    h1 = rnd(gelu(z1))
    z2 = h1 @ rnd(w2).T + b2.astype(np.float64)
    h2 = rnd(gelu(z2))
    r = h2 @ rnd(w3) + float(b3)
    err = r - t
    loss = np.mean(err ** 2)
    d3 = 2.0 * err / n
    g = np.zeros(P)
    g[W3:B3] = h2.T @ d3
    g[B3] = d3.sum()
    d2 = (d3[:, None] * rnd(w3)[None, :]) * gelu_grad(z2)
    g[W2:B2] = (d2.T @ h1).ravel()
    g[B2:W3] = d2.sum(0)
    d1 = (d2 @ rnd(w2)) * gelu_grad(z1)
    g[W1:B1] = (d1.T @ x).ravel()
    g[B1:W2] = d1.sum(0)
    return loss, g


def integer_lifting(q, level, lmin, update=True, inverse=False):
    """Integer interpolet (+ update) transform of an int array stored with
    2^l or 2^l+1 points per axis, in the learned transform's order."""
    a = q.astype(np.int64).copy()
    D = a.ndim
    axes = range(D - 1, -1, -1) if inverse else range(D)
    for d in axes:
        a = np.moveaxis(a, d, -1)
        L = level[d]
        n = 1 << L
        levels = range(lmin[d] + 1, L + 1) if inverse else range(L, lmin[d], -1)
        for lev in levels:
            s = 1 << (L - lev)
            odd = np.arange(s, n, 2 * s)
            even = np.arange(0, n, 2 * s)

            def do_update(sign):
                lo = a[..., (even - s) % n]
                hi = a[..., (even + s) % n]
                a[..., even] += sign * ((lo + hi + 2) >> 2)

            if inverse and update:
                do_update(-1)
            hat = (a[..., (odd - s) % n] + a[..., (odd + s) % n] + 1) >> 1
            a[..., odd] += hat if inverse else -hat
            if update and not inverse:
                do_update(+1)
        if a.shape[-1] == n + 1:
            a[..., n] = a[..., 0]
        a = np.moveaxis(a, -1, d)
    return a
