"""Partial buffers on the GPU (the checked branch of the reference's
lifting pass, paliwa/paliwa_transform.hpp:24-51 with CheckBounds = true)
against the oracle's restatement, bit for bit: sparse coordinate lists
(DDC's SparseDiscreteDomain), affine sub-boxes (DiscreteDomain /
StridedDiscreteDomain), both kernel forms (lines staged in LDS, one launch
per pass for long lines), every wavelet and direction, and the property of
test/test_paliwa_restrict.cpp:94-139."""
import numpy as np
import pytest
import torch

import oracle_bridge as ob
from test_oracle_checked import dependency_closure, lattice

pytestmark = pytest.mark.gpu

WAVELETS = ["hat", "biorthogonal", "fullweighting"]


@pytest.fixture(scope="module")
def pw():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    import paliwa_amd

    return paliwa_amd


def to_storage(pw, coords, kinds):
    """Per axis: Affine for an arithmetic run marked 'affine', else a device
    coordinate list."""
    out = []
    for c, k in zip(coords, kinds):
        if k == "affine":
            step = c[1] - c[0] if len(c) > 1 else 1
            assert c == [c[0] + i * step for i in range(len(c))]
            out.append(pw.Affine(c[0], step))
        else:
            out.append(torch.tensor(c, dtype=torch.int64, device="cuda"))
    return out


def gpu_checked(pw, a, coords, kinds, level, lmin, maximum, name, forward):
    t = torch.from_numpy(a).cuda()
    fn = pw.hierarchize_checked if forward else pw.dehierarchize_checked
    fn(t, to_storage(pw, coords, kinds), level, lmin, maximum, name)
    torch.cuda.synchronize()
    return t.cpu().numpy()


def assert_same(got, want, what):
    if not np.array_equal(got, want):
        diff = np.argwhere(got != want)
        i = tuple(diff[0])
        raise AssertionError(f"{what}: {len(diff)} of {got.size} values differ; first at {i}: "
                             f"gpu {got[i]!r} oracle {want[i]!r}")


@pytest.mark.parametrize("name", WAVELETS)
@pytest.mark.parametrize("forward", [True, False])
def test_restrict_property(pw, name, forward):
    """test/test_paliwa_restrict.cpp:51-139 on the GPU: a sparse buffer of the
    local points and their dependencies, level 6, maximum level 7, every
    minimum level; equal to the oracle's checked branch and, on the local
    points, to the full transform."""
    rng = np.random.default_rng(31)
    level, maximum = 6, 7
    for lmin in range(level):
        start = int(rng.integers(126))
        local = [x for x in lattice(level, maximum)
                 if start <= x < start + 1 + int(rng.integers(max(1, 125 - start)))]
        if not local:
            continue
        storage = sorted(dependency_closure(local, level, maximum, lmin, name, forward)
                         | set(local))
        values = {x: float(rng.uniform(-1, 1)) for x in lattice(level, maximum)}
        full = np.array([values[x] for x in lattice(level, maximum)])
        ob.transform(full, [level], [lmin], [maximum], name, forward)
        buf = np.array([values[x] for x in storage])
        want = ob.transform_checked(buf.copy(), [storage], [level], [lmin], [maximum], name,
                                    forward)
        got = gpu_checked(pw, buf, [storage], ["list"], [level], [lmin], [maximum], name,
                          forward)
        assert_same(got, want, f"lmin {lmin}")
        for i, x in enumerate(storage):
            if x in local:
                assert got[i] == full[x >> (maximum - level)]


def random_axis(rng, level, maximum, kind, count_cap=None):
    P = 1 << maximum
    if kind == "list":
        keep = rng.random(P) < 0.6
        keep[rng.integers(P)] = True
        c = [int(x) for x in np.flatnonzero(keep)]
        return c[:count_cap] if count_cap else c
    step = int(rng.choice([1, 2, 3, 1 << (maximum - level)]))
    front = int(rng.integers(0, P))
    count = 1 + int(rng.integers(max(1, (P - 1 - front) // step + 1)))
    return [front + i * step for i in range(count)]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_product_buffers_against_oracle(pw, dtype):
    """1-D to 4-D buffers mixing coordinate lists and affine runs on every
    axis (on and off the level lattices), all wavelets, both directions."""
    rng = np.random.default_rng(17 if dtype == np.float64 else 18)
    for trial in range(36):
        ndim = 1 + trial % 4
        level = [int(rng.integers(1, 7 - ndim)) for _ in range(ndim)]  # at most 2^16 points
        maximum = [l + int(rng.integers(0, 3)) for l in level]
        lmin = [int(rng.integers(0, l + 1)) for l in level]
        kinds = [("list", "affine")[int(rng.integers(2))] for _ in range(ndim)]
        coords = [random_axis(rng, level[d], maximum[d], kinds[d]) for d in range(ndim)]
        name = WAVELETS[trial % 3]
        forward = bool(trial % 2)
        a = rng.uniform(-1, 1, size=[len(c) for c in coords]).astype(dtype)
        want = ob.transform_checked(a.copy(), coords, level, lmin, maximum, name, forward)
        got = gpu_checked(pw, a, coords, kinds, level, lmin, maximum, name, forward)
        assert_same(got, want, f"trial {trial}: level {level} max {maximum} kinds {kinds}")


def test_dense_sub_box(pw):
    """A DiscreteDomain block (step 1 in finest coordinates, a front, extent
    below the period) of a grid stored at its maximum level, and the same
    block with the grid at a coarser level: only lattice points are swept."""
    rng = np.random.default_rng(5)
    for level, maximum in (([6, 5, 4], [6, 5, 4]), ([5, 4, 4], [6, 5, 4])):
        coords = [list(range(3, 3 + 40)), list(range(0, 31)), list(range(5, 16))]
        kinds = ["affine"] * 3
        for name in WAVELETS:
            for forward in (True, False):
                a = rng.uniform(-1, 1, size=[len(c) for c in coords])
                want = ob.transform_checked(a.copy(), coords, level, [0, 1, 0], maximum, name,
                                            forward)
                got = gpu_checked(pw, a, coords, kinds, level, [0, 1, 0], maximum, name,
                                  forward)
                assert_same(got, want, f"{name} {forward} level {level}")


def test_complete_lattice_takes_the_unchecked_path(pw):
    """A buffer that is the complete lattice of every axis gives exactly the
    ordinary transform (the reference's unchecked branch)."""
    rng = np.random.default_rng(8)
    level, maximum, lmin = [5, 4, 6], [7, 4, 6], [0, 2, 1]
    a = rng.uniform(-1, 1, size=[1 << l for l in level])
    coords = [lattice(l, m) for l, m in zip(level, maximum)]
    for name in WAVELETS:
        got = gpu_checked(pw, a, coords, ["affine"] * 3, level, lmin, maximum, name, True)
        t = torch.from_numpy(a).cuda()
        pw.hierarchize(t, level, lmin, maximum, name)
        assert_same(got, t.cpu().numpy(), name)
        want = ob.transform_checked(a.copy(), coords, level, lmin, maximum, name, True)
        assert_same(got, want, name)


@pytest.mark.parametrize("kind", ["list", "affine"])
def test_long_lines_one_launch_per_pass(pw, kind):
    """Lines beyond the LDS tile (more than 8192 listed coordinates, or more
    than 64 KiB of values) run one launch per pass."""
    rng = np.random.default_rng(21)
    level, maximum = [14, 2], [14, 3]
    if kind == "list":
        c0 = [int(x) for x in np.flatnonzero(rng.random(1 << 14) < 0.7)]
        assert len(c0) > 8192
    else:
        c0 = list(range(100, 100 + 9000))
    coords = [c0, [0, 2, 4, 6, 7]]
    kinds = [kind, "list"]
    for name in ("hat", "biorthogonal"):
        for forward in (True, False):
            a = rng.uniform(-1, 1, size=[len(c) for c in coords])
            want = ob.transform_checked(a.copy(), coords, level, [3, 0], maximum, name, forward)
            got = gpu_checked(pw, a, coords, kinds, level, [3, 0], maximum, name, forward)
            assert_same(got, want, f"{kind} {name} {forward}")


def test_many_lines_grid_stride(pw):
    """More tiles than workgroups (2048 of up to 64 lines): the tile loop
    strides over the grid, on a strided axis (lanes across lines) and on the
    contiguous one (lanes along a line)."""
    rng = np.random.default_rng(4)
    filters = ob.FORWARD["biorthogonal"]
    long_axis = [int(x) for x in np.flatnonzero(rng.random(1 << 18) < 0.9)]
    cases = [  # (level, maximum, coords, kinds, dim)
        ([4, 18], [4, 18], [list(range(16)), long_axis], ["list", "list"], 0),
        ([18, 4], [18, 5], [list(range(1 << 18)), list(range(0, 32, 2))], ["affine", "list"],
         1),
    ]
    for level, maximum, coords, kinds, dim in cases:
        a = rng.uniform(-1, 1, size=[len(c) for c in coords])
        levels = list(range(level[dim], 0, -1))
        want = ob.transform_in_checked(a.copy(), coords, level, maximum, dim, levels, filters)
        t = torch.from_numpy(a).cuda()
        pw.transform_in_checked(dim, t, to_storage(pw, coords, kinds), level, maximum, levels,
                                filters)
        assert_same(t.cpu().numpy(), want, f"{kinds} dim {dim}")


def test_generic_level_and_filter_lists(pw):
    """transform_in on a partial buffer with the reference's multi-pass
    filters (test/test_paliwa_plan.cpp:244-248 style) and non-monotone level
    lists."""
    rng = np.random.default_rng(12)
    filters = [(1, (-0.5, 1.0, -0.5)), (0, (0.25, 1.0, 0.25)), (1, (0.125, 0.75, 0.125)),
               (0, (-0.25, 2.0, 0.5)), (1, (0.3, 0.4, 0.3))]
    coords = [[int(x) for x in np.flatnonzero(rng.random(64) < 0.8)], list(range(1, 8))]
    level, maximum = [6, 3], [6, 3]
    for levels in ([6, 5, 4, 3], [2, 6, 4], [1, 1, 3]):
        a = rng.uniform(-1, 1, size=[len(c) for c in coords])
        want = ob.transform_in_checked(a.copy(), coords, level, maximum, 0, levels, filters)
        t = torch.from_numpy(a).cuda()
        pw.transform_in_checked(0, t, to_storage(pw, coords, ["list", "affine"]), level, maximum,
                                levels, filters)
        assert_same(t.cpu().numpy(), want, str(levels))


def test_axis_variants_and_errors(pw):
    rng = np.random.default_rng(2)
    coords = [list(range(0, 32, 2)), [1, 3, 4, 9]]
    level, maximum, lmin = [4, 3], [5, 4], [0, 0]
    a = rng.uniform(-1, 1, size=[len(c) for c in coords])
    t = torch.from_numpy(a).cuda()
    st = to_storage(pw, coords, ["affine", "list"])
    pw.hierarchize_in_checked(1, t, st, level, lmin, maximum, "biorthogonal")
    pw.dehierarchize_in_checked(0, t, st, level, lmin, maximum, "fullweighting")
    want = a.copy()
    ob.transform_in_checked(want, coords, level, maximum, 1, [3, 2, 1], ob.FORWARD["biorthogonal"])
    ob.transform_in_checked(want, coords, level, maximum, 0, [1, 2, 3, 4],
                            ob.RECONSTRUCT["fullweighting"])
    assert_same(t.cpu().numpy(), want, "per-axis calls")
    with pytest.raises(pw.UnknownWaveletError):
        pw.hierarchize_checked(t, st, level, lmin, maximum, "haar")
    with pytest.raises(pw.ShapeError):
        pw.hierarchize_checked(t, [pw.Affine(20, 2), st[1]], level, lmin, maximum)
