"""numpy checker of the learned-lifting kernels (test infrastructure).

* mlp / mlp_grad: the network on bf16-rounded weights, inputs and
  activations, fp32 biases and the fp16 gate, with fp64 arithmetic
  elsewhere, and its backpropagation (straight through the roundings), to
  check the GPU's MFMA evaluation and gradient kernel.
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


# The kernels' GELU (paliwa_amd/csrc/learned.hip): z * g(z),
#   g = clip(0.5 + s Q(s^2 - 3/8), 0, 1),  s = fp16(z / 4),
# every step an fp16 multiply-add with one rounding; test_learned_ref.py pins
# it against the tanh form, its exact-arithmetic version against the same
# bound, and the derivative against differences.
GATE_SCALE = 0.25
GATE_SHIFT = 0.375
GATE_Q = np.array([0.8046875, -0.9580078125, 1.541015625, -2.6484375, 2.87109375])
GATE_SATURATION = 3.378  # |z| beyond which g is exactly 0 or 1


def gelu_tanh(z):
    """The tanh form the kernels' GELU approximates."""
    return 0.5 * z * (1.0 + np.tanh(0.7978845608028654 * (z + 0.044715 * z ** 3)))


def _f16(x):
    return np.asarray(x, dtype=np.float64).astype(np.float16)


def _fma16(a, b, c):
    """fp16 fused multiply-add: the product of two fp16 values is exact in
    fp64, and so is its sum with an fp16 addend over these ranges; one
    rounding to fp16."""
    return _f16(np.asarray(a, np.float64) * np.asarray(b, np.float64) + np.asarray(c, np.float64))


def gate16(z):
    """g(z) as the kernels evaluate it (fp16 steps), from fp32 or fp64 z."""
    with np.errstate(over="ignore", invalid="ignore"):
        s = _f16(np.asarray(z, dtype=np.float64) * GATE_SCALE)
        w = _fma16(s, s, -GATE_SHIFT)
        p = _fma16(GATE_Q[4], w, GATE_Q[3])
        for k in (2, 1, 0):
            p = _fma16(p, w, GATE_Q[k])
        g = _fma16(s, p, 0.5)
    return np.clip(g.astype(np.float64), 0.0, 1.0)


def _gate_parts(z):
    z = np.asarray(z, dtype=np.float64)
    s = z * GATE_SCALE
    w = s * s - GATE_SHIFT
    p = np.polyval(GATE_Q[::-1], w)
    return z, s, w, p, 0.5 + s * p


def gelu(z, fp16=True):
    """The kernels' activation; fp16=False: the same polynomial in exact
    (fp64) arithmetic, a smooth function for differentiation."""
    if fp16:
        return np.asarray(z, dtype=np.float64) * gate16(z)
    z, s, w, p, raw = _gate_parts(z)
    return z * np.clip(raw, 0.0, 1.0)


def gelu_grad(z):
    """d/dz of gelu(z, fp16=False)."""
    z, s, w, p, raw = _gate_parts(z)
    g = np.clip(raw, 0.0, 1.0)
    dp = np.polyval((GATE_Q[1:] * np.arange(1, 5))[::-1], w)
    dg = GATE_SCALE * (p + 2.0 * s * s * dp)
    inside = (raw > 0.0) & (raw < 1.0)
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
    z1 = x @ rnd(w1).T + b1.astype(np.float64)
    h1 = rnd(gelu(z1, rounding))
    z2 = h1 @ rnd(w2).T + b2.astype(np.float64)
    h2 = rnd(gelu(z2, rounding))
    return h2 @ rnd(w3) + float(b3)


def mlp_grad(u, t, w, rounding=True):
    """(loss, gradient) of mean |mlp(u) - t|, fp64; with rounding, straight
    through the bf16 roundings (d loss / d r = sign(r - t) / n)."""
    rnd = bf16 if rounding else (lambda v: np.asarray(v, dtype=np.float64))
    w1, b1, w2, b2, w3, b3 = unpack(w)
    n = u.shape[0]
    x = rnd(u)
    z1 = x @ rnd(w1).T + b1.astype(np.float64)
    h1 = rnd(gelu(z1, rounding))
    z2 = h1 @ rnd(w2).T + b2.astype(np.float64)
    h2 = rnd(gelu(z2, rounding))
    r = h2 @ rnd(w3) + float(b3)
    err = r - t
    loss = np.mean(np.abs(err))
    d3 = np.sign(err) / n
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
