"""Batched (combination-technique) transform on the GPU: every grid's result
must be bit-identical to the single-grid transform and to the CPU oracle."""
import math

import numpy as np
import pytest
import torch

import oracle_bridge as ob

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pw():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    import paliwa_amd

    return paliwa_amd


def make_grids(levels, image, dtype, seed):
    rng = np.random.default_rng(seed)
    out = []
    for l in levels:
        ext = [(1 << x) + (1 if image else 0) for x in l]
        a = rng.uniform(-1, 1, size=ext).astype(dtype)
        for d, x in enumerate(l):
            if image:
                src = [slice(None)] * len(l)
                dst = [slice(None)] * len(l)
                src[d], dst[d] = 0, 1 << x
                a[tuple(dst)] = a[tuple(src)]
        out.append(np.ascontiguousarray(a))
    return out


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("image", [True, False])
@pytest.mark.parametrize("boxes", [True, False])
def test_batch_matches_oracle(pw, dtype, image, boxes):
    levels = pw.combination_levels(3, (6, 7), 1)
    lmin, lmax = [1, 0, 1], [5, 5, 5]
    arrays = make_grids(levels, image, dtype, seed=21)
    for name in ("hat", "biorthogonal", "fullweighting"):
        for graphs in (True, False):
            grids = [torch.from_numpy(a).cuda() for a in arrays]
            batch = pw.CombinationBatch(grids, levels, lmin, lmax)
            if not boxes:
                batch.set_boxes(0)
            batch.use_graphs(graphs)
            batch.hierarchize(name)
            torch.cuda.synchronize()
            for a, g, l in zip(arrays, grids, levels):
                want = a.copy()
                ob.hierarchize(want, list(l), lmin, lmax, name)
                assert np.array_equal(g.cpu().numpy(), want), (name, l, graphs)
            batch.dehierarchize(name)
            torch.cuda.synchronize()
            for a, g, l in zip(arrays, grids, levels):
                want = a.copy()
                ob.hierarchize(want, list(l), lmin, lmax, name)
                ob.dehierarchize(want, list(l), lmin, lmax, name)
                assert np.array_equal(g.cpu().numpy(), want), (name, l, graphs)
            batch.close()


def test_batch_six_dimensions(pw):
    """77 six-dimensional grids (|l|_1 in {8, 9}, l_i >= 1), with and without
    images, lmin 1 along two axes: six axis phases of groups, each grid
    bit-identical to the oracle after the forward and after the round trip."""
    levels = pw.combination_levels(6, (8, 9), 1)
    assert len(levels) == 77
    lmin, lmax = [0, 1, 0, 0, 1, 0], [4] * 6
    for image in (True, False):
        state = make_grids(levels, image, np.float64, seed=23)
        grids = [torch.from_numpy(a).cuda() for a in state]
        batch = pw.CombinationBatch(grids, levels, lmin, lmax)
        batch.set_boxes(0)  # six axis phases of groups
        # each wavelet starts from the previous one's round trip (not bit-exact)
        for name in ("hat", "biorthogonal"):
            batch.hierarchize(name)
            torch.cuda.synchronize()
            for k, (g, l) in enumerate(zip(grids, levels)):
                ob.hierarchize(state[k], list(l), lmin, lmax, name)
                assert np.array_equal(g.cpu().numpy(), state[k]), (name, l, image)
            batch.dehierarchize(name)
            torch.cuda.synchronize()
            for k, (g, l) in enumerate(zip(grids, levels)):
                ob.dehierarchize(state[k], list(l), lmin, lmax, name)
                assert np.array_equal(g.cpu().numpy(), state[k]), (name, l, image)
        batch.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_batch_wide_tiles(pw, dtype):
    """Strided axes with lines of 512 and 1024 points, whose tiles take 1024
    threads (float64 at 512, float32 at 512 and 1024): every grid against the
    oracle, both directions, hat and biorthogonal."""
    levels = [(9, 3), (10, 2), (9, 9), (3, 10), (10, 4)]
    lmin, lmax = [0, 0], [10, 10]
    arrays = make_grids(levels, True, dtype, seed=24)
    for name in ("hat", "biorthogonal"):
        grids = [torch.from_numpy(a).cuda() for a in arrays]
        batch = pw.CombinationBatch(grids, levels, lmin, lmax)
        batch.set_boxes(0)  # the per-axis groups' wide tiles
        batch.hierarchize(name)
        torch.cuda.synchronize()
        wants = []
        for a, g, l in zip(arrays, grids, levels):
            want = a.copy()
            ob.hierarchize(want, list(l), lmin, lmax, name)
            wants.append(want)
            assert np.array_equal(g.cpu().numpy(), want), (name, l)
        batch.dehierarchize(name)
        torch.cuda.synchronize()
        for want, g, l in zip(wants, grids, levels):
            ob.dehierarchize(want, list(l), lmin, lmax, name)
            assert np.array_equal(g.cpu().numpy(), want), (name, l)
        batch.close()


@pytest.mark.parametrize("dtype,levels", [
    (np.float64, [(9, 4, 9), (9, 5, 8), (4, 9, 9)]),
    (np.float32, [(10, 3, 10), (10, 4, 8), (3, 10, 10)]),
])
def test_batch_groups_over_many_lines(pw, dtype, levels):
    """A strided group of more than 8192 lines of 512 (float64) or 1024
    (float32) points, whose plan builds its tile table for 1024-thread tiles:
    every grid against the oracle, both directions."""
    lmin, lmax = [0, 0, 0], [10, 10, 10]
    arrays = make_grids(levels, True, dtype, seed=25)
    grids = [torch.from_numpy(a).cuda() for a in arrays]
    batch = pw.CombinationBatch(grids, levels, lmin, lmax)
    batch.set_boxes(0)  # the per-axis groups' tile tables
    batch.hierarchize("biorthogonal")
    torch.cuda.synchronize()
    wants = []
    for a, g, l in zip(arrays, grids, levels):
        want = a.copy()
        ob.hierarchize(want, list(l), lmin, lmax, "biorthogonal")
        wants.append(want)
        assert np.array_equal(g.cpu().numpy(), want), l
    batch.dehierarchize("biorthogonal")
    torch.cuda.synchronize()
    for want, g, l in zip(wants, grids, levels):
        ob.dehierarchize(want, list(l), lmin, lmax, "biorthogonal")
        assert np.array_equal(g.cpu().numpy(), want), l
    batch.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_batch_long_lines(pw, dtype):
    """Lines longer than the fused kernel holds (2^12 and 2^13 float64, 2^13
    float32) go through one batched LDS-line launch per group: every grid
    against the oracle, all wavelets, with and without graphs, both
    directions."""
    levels = [(12, 1), (12, 2), (13, 1), (1, 13), (2, 12), (3, 12), (11, 2)]
    lmin, lmax = [0, 1], [13, 13]
    arrays = make_grids(levels, True, dtype, seed=23)
    for name in ("hat", "biorthogonal", "fullweighting"):
        for graphs in (True, False):
            grids = [torch.from_numpy(a).cuda() for a in arrays]
            batch = pw.CombinationBatch(grids, levels, lmin, lmax)
            batch.set_boxes(0)  # the per-axis long-line groups
            batch.use_graphs(graphs)
            batch.hierarchize(name)
            torch.cuda.synchronize()
            wants = []
            for a, g, l in zip(arrays, grids, levels):
                want = a.copy()
                ob.hierarchize(want, list(l), lmin, lmax, name)
                wants.append(want)
                assert np.array_equal(g.cpu().numpy(), want), (name, l, graphs)
            batch.dehierarchize(name)
            torch.cuda.synchronize()
            for want, g, l in zip(wants, grids, levels):
                ob.dehierarchize(want, list(l), lmin, lmax, name)
                assert np.array_equal(g.cpu().numpy(), want), (name, l, graphs)
            batch.close()


def test_batch_replays_graph_repeatedly(pw):
    """Graph replays interleaved with eager work: every replay must equal the
    eager result."""
    levels = pw.combination_levels(2, (7, 8), 1)
    arrays = make_grids(levels, True, np.float64, seed=22)
    ref = [torch.from_numpy(a).cuda() for a in arrays]
    for g, l in zip(ref, levels):
        pw.hierarchize(g, l, [0, 0], [7, 7], "biorthogonal")
    grids = [torch.from_numpy(a).cuda() for a in arrays]
    batch = pw.CombinationBatch(grids, levels, [0, 0], [7, 7])
    for _ in range(3):
        for g, a in zip(grids, arrays):
            g.copy_(torch.from_numpy(a))
        torch.cuda.synchronize()
        batch.hierarchize("biorthogonal")
        torch.cuda.synchronize()
        for g, r in zip(grids, ref):
            assert torch.equal(g, r)


def test_combination_technique_set_1325_grids(pw):
    """The 4-D set |l|_1 in {13..16}, l_i >= 1: 1325 grids, hat basis,
    float64, periodic images: the batched plan (box schedule) bit-identical to
    one hierarchize per grid and to the CPU oracle on every grid, the
    per-axis groups on a sample, and the round trip."""
    from paliwa_amd import fields

    levels = pw.combination_levels(4, (13, 14, 15, 16), 1)
    assert len(levels) == 1325
    rng = np.random.default_rng(2)
    lmin, lmax = [0] * 4, [13] * 4
    grids, singles = [], []
    for l in levels:
        a = rng.uniform(1.0, 4.0, size=4)
        phi = rng.uniform(0.0, 2 * math.pi, size=4)
        g = fields.cosine_product(l, a, phi, device="cuda")
        grids.append(g)
        singles.append(g.clone())
    points = sum(g.numel() for g in grids)
    assert points == 90963469
    sample = list(range(0, 1325, 97))
    initial = {i: grids[i].cpu().numpy() for i in range(len(grids))}
    batch = pw.CombinationBatch(grids, levels, lmin, lmax)
    assert batch.boxes() == (40960, 64, 256)  # the rule's box schedule
    batch.hierarchize("hat")
    for g, l in zip(singles, levels):
        pw.hierarchize(g, l, lmin, lmax, "hat")
    torch.cuda.synchronize()
    for g, s in zip(grids, singles):
        assert torch.equal(g, s)
    # the per-axis groups on a copy: the same bits
    copies = [torch.from_numpy(initial[i]).cuda() for i in sample]
    per_axis = pw.CombinationBatch(copies, [levels[i] for i in sample], lmin, lmax)
    per_axis.set_boxes(0)
    per_axis.hierarchize("hat")
    torch.cuda.synchronize()
    for c, i in zip(copies, sample):
        assert torch.equal(c, grids[i])
    per_axis.close()
    # every grid against the CPU oracle
    for i in range(len(grids)):
        want = initial[i].copy()
        ob.hierarchize(want, list(levels[i]), lmin, lmax, "hat")
        assert np.array_equal(grids[i].cpu().numpy(), want), levels[i]
    batch.dehierarchize("hat")
    torch.cuda.synchronize()
    for i in sample:
        err = np.abs(grids[i].cpu().numpy() - initial[i]).max()
        assert err <= 1e-14, (levels[i], err)
    batch.close()
