"""The combine oracle (oracle/combine.py) pinned against the reference's own
combination-technique test and against what the mathematics fixes."""
import math
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import combine as oc  # noqa: E402
import oracle_bridge as ob  # noqa: E402
from ct_sets import (REFERENCE_GRID_POINTS, REFERENCE_SETS,  # noqa: E402
                     REFERENCE_SPARSE_POINTS, sine_product_grid, standard_set)


@pytest.mark.parametrize("dim", [2, 3, 4])
def test_reference_sizes(dim):
    M, lmin, levels, coefs = REFERENCE_SETS[dim]
    assert sum(int(np.prod([1 << l for l in lv])) for lv in levels) == REFERENCE_GRID_POINTS[dim]
    subs = oc.counted_subspaces(levels)
    assert int(oc.subspace_offsets(subs)[-1]) == REFERENCE_SPARSE_POINTS[dim]


@pytest.mark.parametrize("dim", [2, 3, 4])
def test_reference_mean_of_combined_grid(dim):
    """test/test_paliwa_combination_technique.cpp:150-340: biorthogonal
    hierarchize of sin-product grids, collect, scatter, dehierarchize at the
    maximum level; the mean is (2/pi)^d within 0.031."""
    M, lmin, levels, coefs = REFERENCE_SETS[dim]
    grids = []
    for lv in levels:
        g = sine_product_grid(lv)
        ob.hierarchize(g, lv, lmin, M, "biorthogonal")
        grids.append(g)
    sparse, subs, offs = oc.gather(grids, levels, coefs)
    full = oc.scatter(sparse, subs, offs, M)
    ob.dehierarchize(full, M, lmin, M, "biorthogonal")
    assert abs(full.mean() - (2 / math.pi) ** dim) <= 0.031


def _surplus_check(dim, n, min_count):
    levels, coefs = standard_set(dim, n)
    M = [max(l[d] for l in levels) for d in range(dim)]
    rng = np.random.default_rng(dim * 10 + n)
    full = rng.standard_normal([1 << m for m in M])
    grids = []
    for lv in levels:
        sl = tuple(slice(0, 1 << m, 1 << (m - l)) for m, l in zip(M, lv))
        g = np.ascontiguousarray(full[sl])
        ob.hierarchize(g, lv, [0] * dim, M, "hat")
        grids.append(g)
    surplus = full.copy()
    ob.hierarchize(surplus, M, [0] * dim, M, "hat")
    sparse, subs, offs = oc.gather(grids, levels, coefs, min_count)
    counts = oc.subspace_counts(levels)
    assert len(subs) == sum(1 for c in counts.values() if c >= min_count)
    for si, k in enumerate(subs):
        alpha = surplus[oc._full_slices(k, M)]
        want = np.zeros_like(alpha)
        for l, c in zip(levels, coefs):
            if all(kd <= ld for kd, ld in zip(k, l)):
                want = want + float(c) * alpha
        got = sparse[offs[si]:offs[si + 1]].reshape(alpha.shape)
        assert np.array_equal(got, want), k
    # a valid combination: coefficients over the grids containing a subspace
    # of the sparse grid sum to one
    for k in subs:
        assert sum(c for l, c in zip(levels, coefs)
                   if all(kd <= ld for kd, ld in zip(k, l))) in (0, 1)


@pytest.mark.parametrize("dim,n,min_count", [(2, 7, 2), (2, 7, 1), (3, 7, 2), (4, 7, 1)])
def test_hat_surplus_identity(dim, n, min_count):
    _surplus_check(dim, n, min_count)


def test_pointwise_brute_force_and_scatter():
    levels, coefs = standard_set(3, 5)
    coefs = [c * 1.5 - 0.25 * i for i, c in enumerate(coefs)]  # generic coefficients
    M = [max(l[d] for l in levels) for d in range(3)]
    rng = np.random.default_rng(3)
    grids = [rng.standard_normal([1 << l for l in lv]) for lv in levels]
    for min_count in (1, 2):
        sparse, subs, offs = oc.gather(grids, levels, coefs, min_count)
        brute = oc.gather_pointwise(grids, levels, coefs, M, min_count)
        full = oc.scatter(sparse, subs, offs, M)
        assert len(brute) == sparse.size
        for i, v in brute.items():
            assert full[i] == v, i
        mask = np.zeros(full.shape, dtype=bool)
        for i in brute:
            mask[i] = True
        assert not full[~mask].any()
        # image planes repeat index 0
        img = oc.scatter(sparse, subs, offs, M, [(1 << m) + 1 for m in M])
        assert np.array_equal(img[:-1, :-1, :-1], full)
        assert np.array_equal(img[-1, :-1, :-1], full[0])
        assert np.array_equal(img[:-1, -1, :], img[:-1, 0, :])


def test_float32_grids_accumulate_in_double():
    levels, coefs = [(2, 1), (1, 2), (1, 1)], [1.0, 1.0, -3.0]
    rng = np.random.default_rng(5)
    grids = [rng.standard_normal([1 << l for l in lv]).astype(np.float32) for lv in levels]
    sparse, subs, offs = oc.gather(grids, levels, coefs, 1)
    assert sparse.dtype == np.float32
    # subspace (0, 0): the three grids' index-0 values
    s = np.float32(0)
    for g, c in zip(grids, coefs):
        s = np.float32(np.float64(s) + c * np.float64(g[0, 0]))
    assert sparse[offs[subs.index((0, 0))]] == s
