"""Box schedule of the batched transform (several consecutive axes of a grid
swept together in LDS): every grid bit-identical to the CPU oracle for every
box capacity, tile width and workgroup size, with tiled boxes (partial last
tiles), whole-inner boxes several outer indices per workgroup, single-axis
segments through the groups' kernels, axes without levels, with and without
periodic images."""
import numpy as np
import pytest
import torch

import oracle_bridge as ob

pytestmark = pytest.mark.gpu

WAVELETS = ("hat", "biorthogonal", "fullweighting")


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


def check_round_trip(pw, arrays, levels, lmin, lmax, config, names=WAVELETS):
    for name in names:
        grids = [torch.from_numpy(a).cuda() for a in arrays]
        batch = pw.CombinationBatch(grids, levels, lmin, lmax)
        batch.set_boxes(*config)
        assert batch.boxes() == (tuple(config) if config[0] else (0, 0, 256))
        batch.hierarchize(name)
        torch.cuda.synchronize()
        wants = []
        for a, g, l in zip(arrays, grids, levels):
            want = a.copy()
            ob.hierarchize(want, list(l), lmin, lmax, name)
            wants.append(want)
            assert np.array_equal(g.cpu().numpy(), want), (name, l, config, "forward")
        batch.dehierarchize(name)
        torch.cuda.synchronize()
        for want, g, l in zip(wants, grids, levels):
            ob.dehierarchize(want, list(l), lmin, lmax, name)
            assert np.array_equal(g.cpu().numpy(), want), (name, l, config, "inverse")
        batch.close()


CONFIGS = [
    (40960, 32, 256),   # 4-D grids of 2^12 points split in a tiled and a whole box
    (40960, 4, 512),    # narrow tiles
    (81920, 64, 256),
    (163840, 16, 512),  # most grids in one box
]


@pytest.mark.parametrize("config", CONFIGS)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_box_schedule_4d_set(pw, config, dtype):
    """285 four-dimensional grids, |l|_1 in {11, 12}, with images: every
    wavelet, both directions, against the oracle."""
    levels = pw.combination_levels(4, (11, 12), 1)
    assert len(levels) == 285
    lmin, lmax = [0] * 4, [9] * 4
    arrays = make_grids(levels, True, dtype, seed=31)
    check_round_trip(pw, arrays, levels, lmin, lmax, config)


@pytest.mark.parametrize("config", [(40960, 8, 256), (81920, 32, 512)])
@pytest.mark.parametrize("image", [True, False])
def test_box_schedule_minimum_levels(pw, config, image):
    """Three-dimensional grids with minimum levels 1, 0, 2: axes whose level
    equals the minimum have nothing to sweep (a box may hold such an axis, a
    segment of only such axes is left out); extents 2^l and 2^l + 1."""
    levels = [(5, 5, 4), (2, 8, 2), (1, 3, 9), (6, 0, 2), (1, 7, 7), (7, 7, 2)]
    lmin, lmax = [1, 0, 2], [9, 9, 9]
    arrays = make_grids(levels, image, np.float64, seed=32)
    check_round_trip(pw, arrays, levels, lmin, lmax, config)


def test_box_schedule_long_axes(pw):
    """Grids with one long axis (2^12, 2^13 points) beside short ones: the long
    axis is a single-axis segment (long-line kernels) or a whole box, the
    short ones a box, in either order."""
    levels = [(13, 1, 1), (1, 1, 13), (12, 2, 1), (1, 12, 2), (2, 2, 12), (11, 3, 1)]
    lmin, lmax = [0, 0, 0], [13, 13, 13]
    arrays = make_grids(levels, True, np.float64, seed=33)
    for config in [(81920, 32, 256), (40960, 16, 512)]:
        check_round_trip(pw, arrays, levels, lmin, lmax, config, names=("hat", "biorthogonal"))


def test_box_schedule_six_dimensions(pw):
    """77 six-dimensional grids (|l|_1 in {8, 9}) with minimum levels 1 along
    two axes: boxes of up to six axes."""
    levels = pw.combination_levels(6, (8, 9), 1)
    lmin, lmax = [0, 1, 0, 0, 1, 0], [4] * 6
    arrays = make_grids(levels, True, np.float64, seed=34)
    check_round_trip(pw, arrays, levels, lmin, lmax, (40960, 16, 256))


def test_box_schedule_graph_replays(pw):
    """The box plan replayed as a graph several times, with eager copies in
    between: every replay equals the per-grid transform."""
    levels = pw.combination_levels(4, (10, 11), 1)
    lmin, lmax = [0] * 4, [8] * 4
    arrays = make_grids(levels, True, np.float64, seed=35)
    ref = [torch.from_numpy(a).cuda() for a in arrays]
    for g, l in zip(ref, levels):
        pw.hierarchize(g, l, lmin, lmax, "biorthogonal")
    grids = [torch.from_numpy(a).cuda() for a in arrays]
    batch = pw.CombinationBatch(grids, levels, lmin, lmax)
    batch.set_boxes(40960, 32, 256)
    for _ in range(3):
        for g, a in zip(grids, arrays):
            g.copy_(torch.from_numpy(a))
        torch.cuda.synchronize()
        batch.hierarchize("biorthogonal")
        torch.cuda.synchronize()
        for g, r in zip(grids, ref):
            assert torch.equal(g, r)
    batch.close()


def test_box_settings_are_checked(pw):
    levels = [(3, 3), (2, 4)]
    grids = [torch.zeros((9, 9), dtype=torch.float64, device="cuda"),
             torch.zeros((5, 17), dtype=torch.float64, device="cuda")]
    batch = pw.CombinationBatch(grids, levels, [0, 0], [4, 4])
    before = batch.boxes()
    for bad in [(12345, 32, 256), (81920, 24, 256), (81920, 32, 128), (81920, 0, 256)]:
        with pytest.raises(ValueError):
            batch.set_boxes(*bad)
        assert batch.boxes() == before
    batch.set_boxes(0)
    assert batch.boxes() == (0, 0, 256)
    batch.close()


def test_box_schedule_traffic(pw):
    """The plan's traffic model: 17^4 grids (levels 4, images) take axis 0
    alone and axes 1..3 as one whole box in 40 KiB, so a direction reads
    every point twice instead of four times; the groups read it four times."""
    grids = [torch.zeros((17,) * 4, dtype=torch.float64, device="cuda") for _ in range(3)]
    batch = pw.CombinationBatch(grids, [(4,) * 4] * 3, [0] * 4, [4] * 4)
    batch.set_boxes(40960, 64, 256)
    points = 3 * 17 ** 4
    assert batch.info()["points_read_per_direction"] == 2 * points
    assert batch.info()["launches_per_direction"] == 2
    batch.set_boxes(0)
    assert batch.info()["points_read_per_direction"] == 4 * points
    batch.close()
