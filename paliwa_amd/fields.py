"""Synthetic input fields of the benchmark workloads.

Coordinates follow the reference's periodic unit interval x_i = i / 2^L
(paliwa/paliwa_dimensions.hpp:71-79).  A grid stored with 2^L + 1 points per
axis holds the periodic image x = 1 at index 2^L; every generator closes the
images exactly (index 2^L is a copy of index 0).  Random numbers come from
numpy's PCG64 with the stated seed, so the fields are identical on every
machine; the arithmetic runs in torch on the requested device.
"""
from __future__ import annotations

import math
from typing import Sequence

import numpy as np
import torch


def extents_of(level: Sequence[int], image: bool = True):
    return [(1 << l) + (1 if image else 0) for l in level]


def coords(L: int, extent: int, device, dtype=torch.float64):
    return torch.arange(extent, device=device, dtype=dtype) / float(1 << L)


def close_images(t: torch.Tensor, level: Sequence[int]) -> torch.Tensor:
    for d, l in enumerate(level):
        if t.shape[d] == (1 << l) + 1:
            t.select(d, 1 << l).copy_(t.select(d, 0))
    return t


def _noise(shape, seed: int, device, amplitude: float) -> torch.Tensor:
    rng = np.random.default_rng(seed)
    u = rng.uniform(-1.0, 1.0, size=int(np.prod(shape))).reshape(shape)
    return torch.from_numpy(u).to(device) * amplitude


def gaussian_bump_2d(level=(10, 10), image=True, seed=0, device="cpu"):
    """exp(-20((x-.5)^2+(y-.5)^2)) + 1e-3 U(-1,1): the hat-basis 2-D workload
    (1025 x 1025 with images at level (10, 10))."""
    ext = extents_of(level, image)
    x = coords(level[0], ext[0], device)[:, None]
    y = coords(level[1], ext[1], device)[None, :]
    f = torch.exp(-20.0 * ((x - 0.5) ** 2 + (y - 0.5) ** 2))
    f = f + _noise(ext, seed, device, 1e-3)
    return close_images(f.contiguous(), level)


def fourier_modes(level=(8, 8, 8), modes=16, kmax=32, seed=1, image=True,
                  device="cpu", fronts=0, front_width=0.01):
    """Sum of `modes` Fourier modes cos(2 pi k.x + phi) with integer wave
    vectors k in [1, kmax]^D, amplitude |k|^(-5/3), uniform phases; optional
    tanh fronts across random axis-aligned planes (turbulent-like fields)."""
    rng = np.random.default_rng(seed)
    D = len(level)
    ext = extents_of(level, image)
    xs = [coords(level[d], ext[d], device) for d in range(D)]
    f = torch.zeros(ext, dtype=torch.float64, device=device)
    for _ in range(modes):
        k = rng.integers(1, kmax + 1, size=D)
        phi = rng.uniform(0.0, 2.0 * math.pi)
        amp = float(np.linalg.norm(k)) ** (-5.0 / 3.0)
        # cos(a + b + c) = Re(e^{ia} e^{ib} e^{ic}): separable per axis
        acc = None
        for d in range(D):
            shape = [1] * D
            shape[d] = ext[d]
            e = torch.polar(torch.ones_like(xs[d]), 2.0 * math.pi * float(k[d]) * xs[d])
            e = e.reshape(shape)
            acc = e if acc is None else acc * e
        f += amp * (acc * complex(math.cos(phi), math.sin(phi))).real
    for _ in range(fronts):
        axis = int(rng.integers(0, D))
        pos = rng.uniform(0.0, 1.0)
        height = rng.uniform(0.5, 1.0)
        shape = [1] * D
        shape[axis] = ext[axis]
        # periodic distance to the plane so the front wraps consistently
        dx = torch.remainder(xs[axis] - pos + 0.5, 1.0) - 0.5
        f += height * torch.tanh(dx / front_width).reshape(shape)
    return close_images(f.contiguous(), level)


def cosine_product(level, a, phi, image=True, device="cpu"):
    """prod_i cos(2 pi a_i x_i + phi_i): a component grid of the
    combination-technique workload."""
    ext = extents_of(level, image)
    D = len(level)
    f = None
    for d in range(D):
        shape = [1] * D
        shape[d] = ext[d]
        c = torch.cos(2.0 * math.pi * a[d] * coords(level[d], ext[d], device) + phi[d])
        c = c.reshape(shape)
        f = c if f is None else f * c
    return close_images(f.contiguous(), level)


def cosine_modes_noise(level=(5, 5, 4, 4, 3, 3), seed=3, image=True, device="cpu"):
    """prod_i cos(pi (i+1) x_i) + 1e-3 U(-1,1): the anisotropic 6-D workload."""
    ext = extents_of(level, image)
    D = len(level)
    f = None
    for d in range(D):
        shape = [1] * D
        shape[d] = ext[d]
        c = torch.cos(math.pi * (d + 1) * coords(level[d], ext[d], device)).reshape(shape)
        f = c if f is None else f * c
    f = f + _noise(ext, seed, device, 1e-3)
    return close_images(f.contiguous(), level)
