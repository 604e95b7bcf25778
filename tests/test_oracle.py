"""Pins the CPU oracle (oracle/paliwa_oracle.hpp) before it is trusted as the
checker of the HIP library.

Each check is chosen so that a plausible slip in the restatement -- a wrong
level order, a swapped parity, a neighbour index off by one stride, a dropped
term, a transposed axis -- fails at least one of them:

* the reference's own value known-answer test (test/test_paliwa_plan.cpp:
  114-197), bit-exact, from tests/golden/plan_multipass.json;
* closed forms: the hierarchical surplus of the hat basis, the LeGall 5/3
  (CDF 5/3) analysis filter bank for "biorthogonal", the full-weighting
  stencil, a single Fourier mode, constant fields;
* the reference's invariants: round trip, the hat transform leaving the
  minimum-level lattice untouched and changing every other point
  (test/test_paliwa_transform.cpp:57-64), mass conservation
  (test/test_paliwa_transform.cpp:127-131);
* tensor structure: the D-dimensional transform matrix is the Kronecker
  product of the 1-D ones (brute force from unit vectors);
* the three CPU schedules give bit-identical results.
"""
import json
import os

import numpy as np
import pytest

import oracle_bridge as ob

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
WAVELETS = ["hat", "biorthogonal", "fullweighting"]


def closed_images(a, level):
    """Make index 2^l equal index 0 along every axis stored with an image."""
    for d, l in enumerate(level):
        if a.shape[d] == (1 << l) + 1:
            src = [slice(None)] * a.ndim
            dst = [slice(None)] * a.ndim
            src[d] = 0
            dst[d] = 1 << l
            a[tuple(dst)] = a[tuple(src)]
    return a


def rand_grid(rng, extents, level, dtype=np.float64):
    a = rng.uniform(-1.0, 1.0, size=extents).astype(dtype)
    return np.ascontiguousarray(closed_images(a, level))


# --- the reference's known-answer test ------------------------------------------


def test_reference_multipass_kat_bit_exact():
    cases = json.load(open(os.path.join(GOLDEN, "plan_multipass.json")))
    assert len(cases) == 2 * (2 + 4 + 8 + 16 + 32)
    for c in cases:
        a = np.array([float.fromhex(h) for h in c["input"]])
        filters = [(o, tuple(f)) for o, f in c["filters"]]
        ob.transform_in(a, [c["level"]], 0, c["levels"], filters)
        expected = np.array([float.fromhex(h) for h in c["expected"]])
        # the reference test asserts 1e-12; the restatement is bit-exact
        assert np.array_equal(a, expected), (c["passes"], c["mask"], c["levels"])


# --- closed forms ---------------------------------------------------------------


def hat_surplus_1d(f):
    """Hierarchical surplus of the periodic hat basis from the original nodal
    values: at a point of level l (storage stride s), f(x) - (f(x-s)+f(x+s))/2."""
    n = f.shape[0]
    out = f.copy()
    s = 1
    while s < n:
        x = np.arange(s, n, 2 * s)
        out[x] = (-0.5 * f[(x - s) % n] + f[x]) + (-0.5 * f[(x + s) % n])
        s *= 2
    return out


def test_hat_equals_hierarchical_surplus():
    rng = np.random.default_rng(1)
    for L in range(1, 9):
        f = rng.standard_normal(1 << L)
        a = f.copy()
        ob.hierarchize(a, [L], [0], [L], "hat")
        assert np.array_equal(a, hat_surplus_1d(f)), L


def test_hat_surplus_tensor_2d_3d():
    rng = np.random.default_rng(2)
    for level in ([3, 5], [4, 2, 3]):
        f = rng.standard_normal([1 << l for l in level])
        expect = f.copy()
        for d in range(f.ndim):
            expect = np.apply_along_axis(hat_surplus_1d, d, expect)
        a = np.ascontiguousarray(f.copy())
        ob.hierarchize(a, level, [0] * len(level), level, "hat")
        np.testing.assert_allclose(a, expect, rtol=0, atol=1e-13)


def legall53_analysis(x):
    """One level of the textbook LeGall 5/3 (CDF 5/3) periodic analysis filter
    bank: low-pass taps (-1/8, 1/4, 3/4, 1/4, -1/8) at even samples, high-pass
    taps (-1/2, 1, -1/2) at odd samples."""
    n = x.shape[0]
    e = np.arange(0, n, 2)
    o = np.arange(1, n, 2)
    low = (-x[(e - 2) % n] / 8 + x[(e - 1) % n] / 4 + 3 * x[e] / 4
           + x[(e + 1) % n] / 4 - x[(e + 2) % n] / 8)
    high = x[o] - (x[(o - 1) % n] + x[(o + 1) % n]) / 2
    return low, high


def test_biorthogonal_is_legall53_filter_bank():
    rng = np.random.default_rng(3)
    L = 7
    f = rng.standard_normal(1 << L)
    a = f.copy()
    lmin = 2
    ob.hierarchize(a, [L], [lmin], [L], "biorthogonal")
    coarse = f.copy()
    s = 1
    for _ in range(L - lmin):
        low, high = legall53_analysis(coarse)
        np.testing.assert_allclose(a[s::2 * s], high, rtol=0, atol=1e-13)
        coarse = low
        s *= 2
    np.testing.assert_allclose(a[::s], coarse, rtol=0, atol=1e-13)


def test_fullweighting_closed_form():
    rng = np.random.default_rng(4)
    L = 6
    f = rng.standard_normal(1 << L)
    a = f.copy()
    ob.hierarchize(a, [L], [L - 1], [L], "fullweighting")
    n = f.shape[0]
    e = np.arange(0, n, 2)
    o = np.arange(1, n, 2)
    restricted = f[(e - 1) % n] / 4 + f[e] / 2 + f[(e + 1) % n] / 4
    np.testing.assert_allclose(a[e], restricted, rtol=0, atol=1e-14)
    rn = np.roll(restricted, -1)
    np.testing.assert_allclose(a[o], f[o] - (restricted + rn) / 2, rtol=0, atol=1e-14)


def test_single_fourier_mode_hat_details():
    L, k = 8, 5
    n = 1 << L
    x = np.arange(n)
    f = np.cos(2 * np.pi * k * x / n)
    a = f.copy()
    ob.hierarchize(a, [L], [0], [L], "hat")
    s = 1
    while s < n:
        pts = np.arange(s, n, 2 * s)
        expect = f[pts] * (1 - np.cos(2 * np.pi * k * s / n))
        np.testing.assert_allclose(a[pts], expect, rtol=0, atol=1e-13)
        s *= 2


@pytest.mark.parametrize("name", WAVELETS)
def test_constant_field(name):
    a = np.full((16, 9), 3.25)
    ob.hierarchize(a, [4, 3], [0, 0], [4, 3], name)
    assert a[0, 0] == 3.25
    a[0, 0] = 0.0
    # every detail of a constant vanishes exactly; coarse points keep 3.25
    # for hat/biorthogonal (update weights sum to 1) and fullweighting
    nonzero = a[a != 0.0]
    assert np.all(nonzero == 3.25)


# --- invariants of the reference's tests -----------------------------------------


@pytest.mark.parametrize("name", WAVELETS)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_round_trip(name, dtype):
    rng = np.random.default_rng(5)
    for level, extents, lmin in (
        ([4, 5], [17, 33], [0, 1]),
        ([3, 2, 4], [8, 5, 16], [1, 0, 2]),
        ([2, 1, 3, 2, 1, 2], [5, 3, 9, 4, 2, 5], [0, 0, 1, 0, 0, 1]),
    ):
        a = rand_grid(rng, extents, level, dtype)
        b = a.copy()
        ob.hierarchize(a, level, lmin, level, name)
        ob.dehierarchize(a, level, lmin, level, name)
        tol = 1e-14 if dtype == np.float64 else 3e-6
        assert np.abs(a - b).max() <= tol * max(1.0, np.abs(b).max())


def test_hat_keeps_lmin_lattice_and_changes_the_rest():
    # test/test_paliwa_transform.cpp:17-79: level (3,3) in max (4,4),
    # data (1+ix+iy)(1+U)
    rng = np.random.default_rng(12345)
    for lx in (0, 1, 3):
        for ly in (0, 1, 3):
            ix, iy = np.meshgrid(np.arange(8) * 2, np.arange(8) * 2, indexing="ij")
            f = (1 + ix + iy) * (1 + rng.uniform(0, 1, size=(8, 8)))
            a = np.ascontiguousarray(f.copy())
            ob.hierarchize(a, [3, 3], [lx, ly], [4, 4], "hat")
            on = np.zeros((8, 8), bool)
            on[:: 1 << (3 - lx), :: 1 << (3 - ly)] = True
            np.testing.assert_allclose(a[on], f[on], rtol=0, atol=1e-14)
            assert np.all(a[~on] != f[~on])
            ob.dehierarchize(a, [3, 3], [lx, ly], [4, 4], "hat")
            np.testing.assert_allclose(a, f, rtol=0, atol=1e-13)


@pytest.mark.parametrize("name", ["biorthogonal", "fullweighting"])
def test_mass_conservation(name):
    # test/test_paliwa_transform.cpp:81-143
    rng = np.random.default_rng(12345)
    for lx in (0, 1, 3):
        for ly in (0, 1, 3):
            ix, iy = np.meshgrid(np.arange(8) * 2, np.arange(8) * 2, indexing="ij")
            f = (1 + ix + iy) * (1 + rng.uniform(0, 1, size=(8, 8)))
            a = np.ascontiguousarray(f.copy())
            ob.hierarchize(a, [3, 3], [lx, ly], [4, 4], name)
            lattice = a[:: 1 << (3 - lx), :: 1 << (3 - ly)]
            assert abs(lattice.sum() - f.mean() * lattice.size) <= 1e-13 * abs(f.sum())
            if lx == 0 and ly == 0:
                assert abs(a[0, 0] - f.mean()) <= 1e-14 * abs(f.mean())


# --- tensor structure -------------------------------------------------------------


def transform_matrix(shape, level, lmin, name, forward=True):
    size = int(np.prod(shape))
    M = np.zeros((size, size))
    for j in range(size):
        e = np.zeros(size)
        e[j] = 1.0
        a = np.ascontiguousarray(e.reshape(shape))
        ob.transform(a, level, lmin, level, name, forward)
        M[:, j] = a.ravel()
    return M


@pytest.mark.parametrize("name", WAVELETS)
def test_kronecker_structure(name):
    M = transform_matrix((8, 4), [3, 2], [0, 1], name)
    M0 = transform_matrix((8,), [3], [0], name)
    M1 = transform_matrix((4,), [2], [1], name)
    np.testing.assert_allclose(M, np.kron(M0, M1), rtol=0, atol=1e-15)
    Minv = transform_matrix((8, 4), [3, 2], [0, 1], name, forward=False)
    np.testing.assert_allclose(Minv @ M, np.eye(32), rtol=0, atol=1e-14)


# --- schedules, image plane, golden regression, errors -----------------------------


@pytest.mark.parametrize("name", WAVELETS)
def test_schedules_bit_identical(name):
    rng = np.random.default_rng(6)
    level, extents, lmin = [5, 3, 4], [33, 8, 17], [1, 0, 0]
    base = rand_grid(rng, extents, level)
    outs = []
    for sched in (ob.SERIAL, ob.OMP_DIRECT, ob.OMP_PACKED_POLES):
        a = base.copy()
        ob.hierarchize(a, level, lmin, level, name, sched)
        ob.transform_in(a, level, 1, [1, 3, 2], ob.FORWARD[name], sched)
        outs.append(a)
    assert np.array_equal(outs[0], outs[1])
    assert np.array_equal(outs[0], outs[2])


def test_image_plane_is_lattice_plus_copy():
    rng = np.random.default_rng(7)
    level = [4, 3]
    lattice = rng.standard_normal((16, 8))
    stored = np.zeros((17, 9))
    stored[:16, :8] = lattice
    closed_images(stored, level)
    ob.hierarchize(lattice, level, [0, 0], level, "biorthogonal")
    ob.hierarchize(stored, level, [0, 0], level, "biorthogonal")
    assert np.array_equal(stored[:16, :8], lattice)
    assert np.array_equal(stored[16, :8], lattice[0])
    assert np.array_equal(stored[:16, 8], lattice[:, 0])
    assert stored[16, 8] == lattice[0, 0]


def test_golden_regression_vectors():
    manifest = json.load(open(os.path.join(GOLDEN, "regression.json")))
    for m in manifest:
        dt = np.dtype(m["dtype"])
        load = lambda tag: np.fromfile(
            os.path.join(GOLDEN, f"regression_{m['name']}_{tag}.raw"), dtype=dt
        ).reshape(m["extents"])
        a = np.ascontiguousarray(load("input"))
        ob.hierarchize(a, m["level"], m["minimum_level"], m["level"], m["wavelet"])
        assert np.array_equal(a, load("hierarchized")), m["name"]
        ob.dehierarchize(a, m["level"], m["minimum_level"], m["level"], m["wavelet"])
        assert np.array_equal(a, load("dehierarchized")), m["name"]


def test_oracle_errors():
    a = np.zeros((4, 4))
    with pytest.raises(ob.OracleError) as e:
        ob.hierarchize(a, [2, 2], [0, 0], [2, 2], "haar")
    assert e.value.code == -2
    with pytest.raises(ob.OracleError) as e:
        ob.transform_in(a, [2, 2], 0, [1], [(2, (0.0, 1.0, 0.0))])
    assert e.value.code == -3
    with pytest.raises(ob.OracleError):
        ob.hierarchize(a, [2, 2], [3, 0], [2, 2], "hat")
    with pytest.raises(ob.OracleError):
        ob.hierarchize(np.zeros((5, 3)), [2, 2], [0, 0], [2, 2], "hat")
