"""The oracle's checked branch (paliwa/paliwa_transform.hpp:24-51 with
CheckBounds = true, taken for every buffer that is not the complete strided
lattice) pinned three ways:

* on the complete lattice it equals the unchecked transform (the reference
  takes the unchecked branch there, :224-228, and both must agree);
* a brute force over a dictionary {coordinates: value} of the stored points,
  written from the reference's rule without the oracle's array layout;
* the property of test/test_paliwa_restrict.cpp:94-139: a buffer holding the
  local points plus every input the local outputs depend on reproduces the
  full transform on the local points, bit for bit, while the full lattice
  holds NaN everywhere else.
"""
import itertools

import numpy as np
import pytest

import oracle_bridge as ob

WAVELETS = ("hat", "biorthogonal", "fullweighting")


def lattice(level, maximum):
    return [k << (maximum - level) for k in range(1 << level)]


@pytest.mark.parametrize("name", WAVELETS)
@pytest.mark.parametrize("forward", [True, False])
def test_complete_lattice_equals_unchecked(name, forward):
    rng = np.random.default_rng(7)
    level, maximum, lmin = [4, 3, 5], [5, 3, 6], [0, 1, 2]
    a = rng.uniform(-1, 1, size=[1 << l for l in level])
    want = a.copy()
    ob.transform(want, level, lmin, level, name, forward)
    got = a.copy()
    ob.transform_checked(got, [lattice(l, m) for l, m in zip(level, maximum)], level, lmin,
                         maximum, name, forward)
    assert np.array_equal(got, want)


def brute_force(values, coords, level, maximum, dim, levels, filters, dtype):
    """Dictionary model: every stored point by its coordinates; a pass
    updates a point of its half lattice whose two neighbours are stored."""
    pts = {}
    for idx in itertools.product(*[range(len(c)) for c in coords]):
        pts[tuple(coords[d][i] for d, i in enumerate(idx))] = dtype(values[idx])
    P = 1 << maximum[dim]
    for l in levels:
        t = 1 << (maximum[dim] - l)
        for offset, f in filters:
            c0, c1, c2 = (dtype(v) for v in f)
            new = {}
            for x, v in pts.items():
                if x[dim] % (2 * t) != offset * t:
                    continue
                if any(x[d] % (1 << (maximum[d] - level[d])) for d in range(len(x)) if d != dim):
                    continue
                lo = list(x)
                hi = list(x)
                lo[dim] = P - t if x[dim] == 0 else x[dim] - t
                hi[dim] = 0 if x[dim] + t == P else x[dim] + t
                lo, hi = tuple(lo), tuple(hi)
                if lo in pts and hi in pts:
                    new[x] = dtype(dtype(dtype(c0 * pts[lo]) + dtype(c1 * v)) + dtype(c2 * pts[hi]))
            pts.update(new)
    out = np.empty_like(values)
    for idx in itertools.product(*[range(len(c)) for c in coords]):
        out[idx] = pts[tuple(coords[d][i] for d, i in enumerate(idx))]
    return out


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_brute_force_on_random_buffers(dtype):
    rng = np.random.default_rng(11)
    for trial in range(40):
        ndim = 1 + trial % 3
        level = [int(rng.integers(1, 4)) for _ in range(ndim)]
        maximum = [l + int(rng.integers(0, 2)) for l in level]
        coords = []
        for d in range(ndim):
            P = 1 << maximum[d]
            keep = rng.random(P) < 0.7
            keep[rng.integers(P)] = True
            coords.append([int(x) for x in np.flatnonzero(keep)])
        dim = int(rng.integers(ndim))
        levels = list(range(level[dim], 0, -1)) if rng.integers(2) else list(range(1, level[dim] + 1))
        filters = [(int(rng.integers(2)), tuple(float(v) for v in rng.choice(
            [-0.5, -0.25, 0.25, 0.5, 1.0, 2.0, 0.3], 3))) for _ in range(1 + trial % 3)]
        a = rng.uniform(-1, 1, size=[len(c) for c in coords]).astype(dtype)
        want = brute_force(a, coords, level, maximum, dim, levels, filters, dtype)
        got = ob.transform_in_checked(a.copy(), coords, level, maximum, dim, levels, filters)
        assert np.array_equal(got, want), trial


def dependency_closure(outputs, level, maximum, lmin, name, forward):
    """Every input the given output coordinates depend on: the passes walked
    backwards, each write making its two neighbours required (the stencil
    graph of test/test_paliwa_plan.cpp:34-69)."""
    P = 1 << maximum
    filters = (ob.FORWARD if forward else ob.RECONSTRUCT)[name]
    levels = list(range(level, lmin, -1)) if forward else list(range(lmin + 1, level + 1))
    req = set(outputs)
    for l in reversed(levels):
        t = 1 << (maximum - l)
        for offset, _ in reversed(filters):
            writes = [x for x in req if x % (2 * t) == offset * t]
            for x in writes:
                req.add((x - t) % P)
                req.add((x + t) % P)
    return req


@pytest.mark.parametrize("name", WAVELETS)
@pytest.mark.parametrize("forward", [True, False])
def test_restrict_property_1d(name, forward):
    """test/test_paliwa_restrict.cpp:51-139: level 6, maximum level 7, every
    minimum level, a local run of random start and extent (seeded here)."""
    rng = np.random.default_rng(2025)
    level, maximum = 6, 7
    full_coords = lattice(level, maximum)
    for lmin in range(level):
        start = int(rng.integers(126))
        extent = 1 + int(rng.integers(max(1, 125 - start)))
        local = [x for x in full_coords if start <= x < start + extent]
        if not local:
            continue
        storage = sorted(dependency_closure(local, level, maximum, lmin, name, forward) | set(local))
        full = np.full(1 << level, np.nan)
        buf = np.empty(len(storage))
        for i, x in enumerate(storage):
            v = 1.0 if x in local else 2.0
            full[x >> (maximum - level)] = v
            buf[i] = v
        ob.transform(full, [level], [lmin], [maximum], name, forward)
        ob.transform_checked(buf, [storage], [level], [lmin], [maximum], name, forward)
        for i, x in enumerate(storage):
            if x in local:
                assert not np.isnan(full[x >> (maximum - level)])
                assert buf[i] == full[x >> (maximum - level)], (lmin, x)


@pytest.mark.parametrize("name", WAVELETS)
def test_restrict_property_2d_random_values(name):
    """The same property on a 2-D product buffer (each axis: a local run plus
    its own dependency closure) with random values: every local point equals
    the full transform bit for bit, both directions."""
    rng = np.random.default_rng(99)
    level, maximum, lmin = [5, 4], [6, 4], [1, 0]
    for forward in (True, False):
        local, coords = [], []
        for d in range(2):
            lat = lattice(level[d], maximum[d])
            s = int(rng.integers(len(lat)))
            run = [lat[(s + k) % len(lat)] for k in range(1 + int(rng.integers(len(lat) // 3)))]
            local.append(set(run))
            coords.append(sorted(dependency_closure(run, level[d], maximum[d], lmin[d], name,
                                                    forward) | set(run)))
        full = rng.uniform(-1, 1, size=[1 << l for l in level])
        buf = np.empty([len(c) for c in coords])
        for i, x in enumerate(coords[0]):
            for j, y in enumerate(coords[1]):
                buf[i, j] = full[x >> (maximum[0] - level[0]), y >> (maximum[1] - level[1])]
        ob.transform(full, level, lmin, maximum, name, forward)
        ob.transform_checked(buf, coords, level, lmin, maximum, name, forward)
        for i, x in enumerate(coords[0]):
            for j, y in enumerate(coords[1]):
                if x in local[0] and y in local[1]:
                    assert buf[i, j] == full[x >> (maximum[0] - level[0]),
                                             y >> (maximum[1] - level[1])]


def test_a_missing_dependency_is_detected():
    """Dropping one required input changes some local output: the property
    is sensitive to the storage set (a planted gap in the oracle's input)."""
    level, maximum, lmin, name = 6, 7, 0, "biorthogonal"
    local = [x for x in lattice(level, maximum) if 40 <= x < 60]
    req = sorted(dependency_closure(local, level, maximum, lmin, name, True) | set(local))
    ghost = [x for x in req if x not in local]
    rng = np.random.default_rng(3)
    vals = {x: float(rng.uniform(-1, 1)) for x in lattice(level, maximum)}
    full = np.array([vals[x] for x in lattice(level, maximum)])
    ob.transform(full, [level], [lmin], [maximum], name, True)
    changed = 0
    for drop in ghost:
        storage = [x for x in req if x != drop]
        buf = np.array([vals[x] for x in storage])
        ob.transform_checked(buf, [storage], [level], [lmin], [maximum], name, True)
        if any(buf[i] != full[x >> 1] for i, x in enumerate(storage) if x in local):
            changed += 1
    assert changed == len(ghost)
