"""Combination-technique combine on the GPU against the combine oracle
(oracle/combine.py): the reference test's three sets, float32, both subspace
rules, periodic images, and the 1325-grid set."""
import math
import os
import sys

import numpy as np
import pytest
import torch

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import combine as oc  # noqa: E402
import oracle_bridge as ob  # noqa: E402
from ct_sets import REFERENCE_SETS, sine_product_grid, standard_set  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pw():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    import paliwa_amd

    return paliwa_amd


def _hierarchized(levels, lmin, M, wavelet, dtype=np.float64):
    grids = []
    for lv in levels:
        g = sine_product_grid(lv, dtype)
        ob.hierarchize(g, lv, lmin, M, wavelet)
        grids.append(g)
    return grids


@pytest.mark.parametrize("dim", [2, 3, 4])
def test_reference_sets_bit_exact(pw, dim):
    """test/test_paliwa_combination_technique.cpp:150-340 end to end: batched
    biorthogonal hierarchize, collect, scatter, dehierarchize; every stage
    bit-identical to the oracle, and the reference's mean check."""
    from paliwa_amd.combine import SparseGridCombination

    M, lmin, levels, coefs = REFERENCE_SETS[dim]
    nodal = [sine_product_grid(lv) for lv in levels]
    grids = [torch.from_numpy(g.copy()).cuda() for g in nodal]
    batch = pw.CombinationBatch(grids, levels, lmin, M)
    batch.hierarchize("biorthogonal")
    want_grids = _hierarchized(levels, lmin, M, "biorthogonal")
    for g, w in zip(grids, want_grids):
        assert np.array_equal(g.cpu().numpy(), w)
    comb = SparseGridCombination(grids, levels, coefs, M)
    want_sparse, subs, offs = oc.gather(want_grids, levels, coefs)
    assert comb.subspaces == subs
    assert comb.offsets == list(offs)
    sparse = comb.gather()
    assert np.array_equal(sparse.cpu().numpy(), want_sparse)
    for image in (False, True):
        shape = [(1 << m) + int(image) for m in M]
        full = comb.scatter(sparse, image=image)
        assert np.array_equal(full.cpu().numpy(), oc.scatter(want_sparse, subs, offs, M, shape))
    full = comb.combine_full(lmin, "biorthogonal")
    want_full = oc.scatter(want_sparse, subs, offs, M)
    ob.dehierarchize(want_full, M, lmin, M, "biorthogonal")
    got = full.cpu().numpy()
    assert np.array_equal(got, want_full)
    assert abs(got.mean() - (2 / math.pi) ** dim) <= 0.031
    comb.close()


@pytest.mark.parametrize("min_count", [1, 2, 3])
def test_subspace_rules_and_float32(pw, min_count):
    from paliwa_amd.combine import SparseGridCombination

    levels, coefs = standard_set(3, 8)
    coefs = [c + 0.125 * (i % 5) for i, c in enumerate(coefs)]  # inexact products
    M = [max(l[d] for l in levels) for d in range(3)]
    rng = np.random.default_rng(min_count)
    for dtype, tdtype in ((np.float64, torch.float64), (np.float32, torch.float32)):
        host = [rng.standard_normal([(1 << l) + (i % 2) for l in lv]).astype(dtype)
                for i, lv in enumerate(levels)]
        grids = [torch.from_numpy(h).cuda() for h in host]
        comb = SparseGridCombination(grids, levels, coefs, M, min_count=min_count)
        want, subs, offs = oc.gather(host, levels, coefs, min_count)
        got = comb.gather()
        assert got.dtype == tdtype
        assert comb.subspaces == subs
        assert np.array_equal(got.cpu().numpy(), want), (dtype, min_count)
        full = comb.scatter(got, image=True)
        assert np.array_equal(full.cpu().numpy(),
                              oc.scatter(want, subs, offs, M, [(1 << m) + 1 for m in M]))
        comb.close()


def test_one_five_and_eight_dimensions(pw):
    from paliwa_amd.combine import SparseGridCombination

    for dim, n in ((1, 1), (1, 9), (5, 8), (8, 10)):
        levels, coefs = standard_set(dim, n, minimum=1 if n > 1 else 0)
        M = [max(l[d] for l in levels) for d in range(dim)]
        rng = np.random.default_rng(dim)
        host = [rng.standard_normal([1 << l for l in lv]) for lv in levels]
        grids = [torch.from_numpy(h).cuda() for h in host]
        for min_count in (1, 2):
            comb = SparseGridCombination(grids, levels, coefs, M, min_count=min_count)
            want, subs, offs = oc.gather(host, levels, coefs, min_count)
            if want.size == 0:
                assert comb.sparse_points == 0
                continue
            assert np.array_equal(comb.gather().cpu().numpy(), want), (dim, n, min_count)
            comb.close()


def test_combination_technique_set_1325_grids(pw):
    """The 4-D set |l|_1 in {13..16}, l_i >= 1, coefficients +1, -3, +3, -1 for
    |l|_1 = 16, 15, 14, 13: hat hierarchize, then the combine, against the
    oracle."""
    from paliwa_amd import fields
    from paliwa_amd.combine import SparseGridCombination

    levels = pw.combination_levels(4, (13, 14, 15, 16), 1)
    coef_of = {16: 1, 15: -3, 14: 3, 13: -1}
    coefs = [coef_of[sum(l)] for l in levels]
    rng = np.random.default_rng(2)
    grids = []
    for l in levels:
        a = rng.uniform(1.0, 4.0, size=4)
        phi = rng.uniform(0.0, 2 * math.pi, size=4)
        grids.append(fields.cosine_product(l, a, phi, device="cuda"))
    M = [13] * 4
    batch = pw.CombinationBatch(grids, levels, [0] * 4, M)
    batch.hierarchize("hat")
    comb = SparseGridCombination(grids, levels, coefs, M)
    info = comb.info()
    assert info["subspaces"] == 3338 and info["sparse_points"] == 2600960
    assert info["terms"] == 44544000
    sparse = comb.gather()
    host = [g.cpu().numpy() for g in grids]
    want, subs, offs = oc.gather(host, levels, coefs)
    assert np.array_equal(sparse.cpu().numpy(), want)
    comb.close()
    batch.close()


def test_wide_records_for_large_grids(pw):
    """A grid of 2^24 elements or more makes the plan keep 64-bit prefix
    multipliers and 64-bit index arithmetic (the narrow 24-bit form holds
    only smaller grids): a 257^3 grid with three of its coarser neighbours,
    against the oracle, both subspace rules."""
    from paliwa_amd.combine import SparseGridCombination

    levels = [(8, 8, 8), (8, 8, 7), (8, 7, 8), (7, 8, 8)]
    coefs = [1.0, -0.375, -0.375, -0.25]
    M = [8, 8, 8]
    rng = np.random.default_rng(7)
    host = [rng.standard_normal([(1 << l) + (1 if i == 0 else 0) for l in lv])
            for i, lv in enumerate(levels)]
    assert host[0].size >= 1 << 24
    grids = [torch.from_numpy(h).cuda() for h in host]
    for min_count in (1, 2):
        comb = SparseGridCombination(grids, levels, coefs, M, min_count=min_count)
        want, subs, offs = oc.gather(host, levels, coefs, min_count)
        assert comb.subspaces == subs
        assert np.array_equal(comb.gather().cpu().numpy(), want), min_count
        comb.close()
