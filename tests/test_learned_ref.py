"""Pins the numpy checker of the learned lifting (tests/learned_ref.py)."""
import numpy as np

import learned_ref as lr


def test_backprop_matches_finite_differences():
    # rounding-free network: smooth, and no residual near the kink of |.|
    # (each differs from 0 by far more than a step moves it), so central
    # differences apply
    rng = np.random.default_rng(0)
    w = (rng.standard_normal(lr.P) * 0.3).astype(np.float32)
    u = rng.uniform(-1, 1, size=(40, 8))
    t = rng.uniform(-0.5, 0.5, size=40)
    assert np.abs(mlp64(u, w.astype(np.float64)) - t).min() > 1e-3
    _, g = lr.mlp_grad(u, t, w, rounding=False)
    w64 = w.astype(np.float64)
    for e in list(range(0, lr.P, 37)) + [lr.B3, lr.W3 + 5, lr.B1 + 3]:
        h = 1e-6
        wp, wm = w64.copy(), w64.copy()
        wp[e] += h
        wm[e] -= h
        fp = np.mean(np.abs(mlp64(u, wp) - t))
        fm = np.mean(np.abs(mlp64(u, wm) - t))
        fd = (fp - fm) / (2 * h)
        assert abs(fd - g[e]) <= 1e-6 + 1e-5 * abs(fd), (e, fd, g[e])


def mlp64(u, w):
    w1, b1 = w[lr.W1:lr.B1].reshape(32, 8), w[lr.B1:lr.W2]
    w2, b2 = w[lr.W2:lr.B2].reshape(32, 32), w[lr.B2:lr.W3]
    w3, b3 = w[lr.W3:lr.B3], w[lr.B3]
    h1 = lr.gelu(u @ w1.T + b1, fp16=False)
    h2 = lr.gelu(h1 @ w2.T + b2, fp16=False)
    return h2 @ w3 + b3


def all_fp16_inputs():
    """z = 4 s for every finite fp16 s (the gate's argument), both signs."""
    s = np.arange(0, 0x7C00, dtype=np.uint16).view(np.float16).astype(np.float64)
    return 4.0 * np.concatenate([-s[::-1], s])


def test_gelu_is_the_tanh_form_within_its_fit():
    """The kernels' transcendental-free GELU: within 2.05e-3 of the tanh form
    for every fp16 argument with its fp16 roundings, within 1.1e-3 in exact
    arithmetic (a dense grid); exactly z and 0 beyond the saturation point,
    no NaN anywhere, the exact gate monotone and odd-symmetric about 1/2
    (gelu(z) - gelu(-z) = z)."""
    z = all_fp16_inputs()
    h = lr.gelu(z)
    assert not np.isnan(h).any()
    small = np.abs(z) < 1e4  # gelu_tanh itself overflows further out
    err = np.abs(h[small] - lr.gelu_tanh(z[small]))
    assert err.max() <= 2.05e-3, err.max()
    sat = np.abs(z) >= lr.GATE_SATURATION + 0.01
    assert np.array_equal(h[sat & (z > 0)], z[sat & (z > 0)])
    assert np.all(h[sat & (z < 0)] == 0.0)
    zz = np.concatenate([np.linspace(-12, 12, 240001), np.linspace(-60, 60, 12001)])
    exact = lr.gelu(zz, fp16=False)
    assert np.abs(exact - lr.gelu_tanh(zz)).max() <= 1.1e-3
    assert np.allclose(exact - lr.gelu(-zz, fp16=False), zz, rtol=0, atol=1e-6)
    g = lr.gelu(np.linspace(0, 12, 120001), fp16=False) / np.linspace(1e-300, 12, 120001)
    assert np.all(np.diff(g[1:]) >= -1e-12)


def test_gate_fp16_steps_round_once():
    """The fp16 restatement's fused multiply-add rounds once: an fp16 product
    that is exact in fp64 plus an addend, against the value rounded twice
    (product to fp16 first) on a case where they differ."""
    a = b = np.float16(1 + 2.0 ** -10)
    c = np.float16(-(1 + 2.0 ** -9))
    once = lr._fma16(a, b, c)  # (1 + 2^-10)^2 - (1 + 2^-9) = 2^-20, an fp16 subnormal
    twice = np.float16(np.float16(a * b) + c)
    assert float(once) == 2.0 ** -20 and float(twice) == 0.0


def test_gelu_grad_matches_differences():
    rng = np.random.default_rng(5)
    z = np.concatenate([rng.uniform(-5, 5, 4000), [0.0, 1e-3, -1e-3, 3.0, -3.0]])
    z = z[np.abs(np.abs(z) - lr.GATE_SATURATION) > 1e-2]  # away from the kink
    h = 1e-6
    fd = (lr.gelu(z + h, fp16=False) - lr.gelu(z - h, fp16=False)) / (2 * h)
    assert np.allclose(lr.gelu_grad(z), fd, rtol=1e-6, atol=1e-7)


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -2.5], dtype=np.float32)
    # 7 stored mantissa bits: 1 + 2^-8 is a tie and goes to even (1),
    # 1 + 2^-9 rounds down, 1 + 3*2^-9 rounds up to 1 + 2^-7
    assert list(lr.bf16(x)) == [1.0, 1.0, 1.0, 1.0 + 2 ** -7, -2.5]


def test_integer_lifting_round_trip_and_hat_closed_form():
    rng = np.random.default_rng(1)
    level, lmin = [4, 3, 5], [0, 1, 2]
    q = rng.integers(-1000, 1000, size=[17, 8, 33])
    q[16] = q[0]
    q[:, :, 32] = q[:, :, 0]
    for update in (False, True):
        f = lr.integer_lifting(q, level, lmin, update)
        b = lr.integer_lifting(f, level, lmin, update, inverse=True)
        assert np.array_equal(b, q)
    # without update, one axis: detail = x - floor((left + right + 1) / 2)
    line = rng.integers(-50, 50, size=16)
    f = lr.integer_lifting(line, [4], [3], update=False)
    odd = np.arange(1, 16, 2)
    assert np.array_equal(f[odd], line[odd] - np.floor((line[odd - 1] + line[(odd + 1) % 16] + 1) / 2))
