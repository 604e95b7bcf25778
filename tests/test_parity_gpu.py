"""Parity of the HIP library with the CPU oracle on the GPU.

Float64 and float32 results must be bit-identical to the oracle's (the
lifting update is evaluated in the reference's operand order with no fused
multiply-add, so there is nothing to round differently).  At the full
workload sizes the checks are the oracle (OpenMP schedule, bit-identical to
the serial one) plus size-independent properties: round trip and mass.
"""
import json
import os

import numpy as np
import pytest
import torch

import oracle_bridge as ob

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
WAVELETS = ["hat", "biorthogonal", "fullweighting"]


@pytest.fixture(scope="module")
def pw():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    import paliwa_amd

    return paliwa_amd


def closed_random(rng, extents, level, dtype):
    a = rng.uniform(-1.0, 1.0, size=extents).astype(dtype)
    for d, l in enumerate(level):
        if extents[d] == (1 << l) + 1:
            src = [slice(None)] * len(extents)
            dst = [slice(None)] * len(extents)
            src[d], dst[d] = 0, 1 << l
            a[tuple(dst)] = a[tuple(src)]
    return np.ascontiguousarray(a)


def gpu_named(pw, a, level, lmin, name, forward):
    t = torch.from_numpy(a).cuda()
    (pw.hierarchize if forward else pw.dehierarchize)(t, level, lmin, level, name)
    torch.cuda.synchronize()
    return t.cpu().numpy()


def oracle_named(a, level, lmin, name, forward, schedule=ob.SERIAL):
    b = a.copy()
    ob.transform(b, level, lmin, level, name, forward, schedule)
    return b


def assert_same(got, want, what):
    if not np.array_equal(got, want):
        diff = np.argwhere(got != want)
        i = tuple(diff[0])
        raise AssertionError(
            f"{what}: {len(diff)} of {got.size} values differ; first at {i}: "
            f"gpu {got[i]!r} oracle {want[i]!r}")


def line_shapes(L):
    n = 1 << L
    return [
        ([2, 1, L], [5, 3, n + 1]),   # contiguous axis, image, 15 lines
        ([2, 1, L], [4, 2, n]),       # contiguous axis, no image
        ([L, 2, 1], [n + 1, 5, 3]),   # strided axis, B = 15
        ([L, 6], [n, 65]),            # strided axis, B = 65
        ([L], [n + 1]),               # one line
    ]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_every_line_length_bit_exact(pw, dtype):
    rng = np.random.default_rng(11)
    maxL = 11 if dtype == np.float64 else 12
    for L in range(1, maxL + 1):
        for level, extents in line_shapes(L):
            for name in WAVELETS:
                for lmin_L in sorted({0, L // 2, L - 1}):
                    lmin = [min(lmin_L, l) for l in level]
                    a = closed_random(rng, extents, level, dtype)
                    for fwd in (True, False):
                        got = gpu_named(pw, a, level, lmin, name, fwd)
                        want = oracle_named(a, level, lmin, name, fwd)
                        assert_same(got, want, f"L={L} {extents} {name} lmin={lmin} fwd={fwd}")


@pytest.mark.parametrize("name", WAVELETS)
def test_multidimensional_bit_exact(pw, name):
    rng = np.random.default_rng(12)
    cases = [
        ([6, 5], [65, 33], [0, 2]),
        ([4, 3, 5], [17, 9, 33], [0, 0, 1]),
        ([3, 3, 2, 2, 1, 1], [9, 9, 5, 5, 3, 3], [0, 1, 0, 0, 0, 1]),
        ([7, 2, 8], [128, 5, 257], [3, 0, 0]),
    ]
    for dtype in (np.float64, np.float32):
        for level, extents, lmin in cases:
            a = closed_random(rng, extents, level, dtype)
            fwd = gpu_named(pw, a, level, lmin, name, True)
            assert_same(fwd, oracle_named(a, level, lmin, name, True), f"{extents} fwd")
            back = gpu_named(pw, fwd, level, lmin, name, False)
            assert_same(back, oracle_named(fwd, level, lmin, name, False), f"{extents} inv")


def test_maximum_dimension_count(pw):
    """26 axes (the C ABI's limit): level-0 axes of one or two stored points,
    level-1 axes with and without the periodic image, one longer axis; every
    axis sweeps a different [A][m][B] view."""
    rng = np.random.default_rng(16)
    spec = [(1, 2), (0, 1), (1, 3), (0, 1), (0, 2), (1, 2), (2, 5), (0, 1), (0, 1), (1, 2),
            (1, 2), (0, 1), (1, 3), (0, 1), (0, 1), (1, 2), (0, 2), (0, 1), (1, 3), (1, 2),
            (0, 1), (1, 2), (0, 1), (0, 1), (0, 1), (3, 8)]  # 552,960 points
    assert len(spec) == 26
    level = [l for l, _ in spec]
    extents = [e for _, e in spec]
    for name in WAVELETS:
        for dtype in (np.float64, np.float32):
            a = closed_random(rng, extents, level, dtype)
            lmin = [0] * 26
            fwd = gpu_named(pw, a, level, lmin, name, True)
            assert_same(fwd, oracle_named(a, level, lmin, name, True), f"26-d {name} fwd")
            back = gpu_named(pw, fwd, level, lmin, name, False)
            assert_same(back, oracle_named(fwd, level, lmin, name, False), f"26-d {name} inv")


@pytest.mark.parametrize("dtype,level,extents", [
    (np.float64, [9, 4, 9], [513, 17, 513]),
    (np.float64, [10, 3, 10], [1025, 9, 1025]),
    (np.float32, [9, 4, 9], [513, 17, 513]),
    (np.float32, [10, 3, 10], [1025, 9, 1025]),
    (np.float32, [11, 3, 10], [2049, 9, 1025]),
])
def test_wide_tiles_over_many_lines(pw, dtype, level, extents):
    """Strided sweeps over more than 8192 lines of 512 to 2048 points, which
    take the 1024-thread ladder (the small-grid tests above run the same
    periods with 512 threads): both directions against the oracle."""
    rng = np.random.default_rng(17)
    for name in ("hat", "biorthogonal"):
        a = closed_random(rng, extents, level, dtype)
        lmin = [0, 0, 0]
        fwd = gpu_named(pw, a, level, lmin, name, True)
        assert_same(fwd, oracle_named(a, level, lmin, name, True), f"{extents} {name} fwd")
        back = gpu_named(pw, fwd, level, lmin, name, False)
        assert_same(back, oracle_named(fwd, level, lmin, name, False), f"{extents} {name} inv")


@pytest.mark.parametrize("path", ["fused", "lds_line", "per_pass"])
def test_kernel_paths_bit_exact(pw, path):
    rng = np.random.default_rng(13)
    for dtype in (np.float64, np.float32):
        for level, extents, dim in (([5, 3], [33, 9], 0), ([3, 7], [8, 129], 1),
                                    ([6, 2, 4], [64, 5, 17], 1)):
            for name in WAVELETS:
                for table in (ob.FORWARD, ob.RECONSTRUCT):
                    levels = list(range(level[dim], 0, -1))
                    if table is ob.RECONSTRUCT:
                        levels = levels[::-1]
                    a = closed_random(rng, extents, level, dtype)
                    t = torch.from_numpy(a).cuda()
                    pw.transform_in(dim, t, level, level, levels, table[name], path=path)
                    want = a.copy()
                    ob.transform_in(want, level, dim, levels, table[name])
                    assert_same(t.cpu().numpy(), want, f"{path} {extents} dim={dim} {name}")


def test_generic_level_and_filter_lists(pw):
    """Non-consecutive, repeated and partial level lists and 3..5-pass custom
    filters (test/test_paliwa_plan.cpp:244-248 uses 5-pass filters)."""
    rng = np.random.default_rng(14)
    filter_sets = [
        [(0, (0.5, 0.25, 0.125)), (1, (-0.5, 1.0, 0.125)), (1, (0.0, 1.0, 0.25))],
        [(1, (-0.5, 1.0, 0.125)), (0, (0.0, 1.0, 0.25)), (1, (-0.5, 1.0, 0.125)),
         (0, (0.0, 1.0, 0.25)), (1, (0.3, 0.7, -0.1))],
        [(0, (0.25, 1.0, 0.25))],
        [(1, (0.1, 1.0, 0.2)), (1, (-0.3, 1.0, 0.4))],
    ]
    level_lists = [[3, 1, 2], [5, 5, 2], [1, 2, 3, 4, 5], [2, 3], [4, 3], [5], [1]]
    for dtype in (np.float64, np.float32):
        for level, extents, dim in (([5, 4], [33, 16], 0), ([4, 5], [17, 33], 1),
                                    ([3, 5, 2], [9, 32, 4], 1)):
            for filters in filter_sets:
                for levels in level_lists:
                    levels = [l for l in levels if l <= level[dim]]
                    if not levels:
                        continue
                    for path in ("auto", "lds_line", "per_pass"):
                        a = closed_random(rng, extents, level, dtype)
                        t = torch.from_numpy(a).cuda()
                        pw.transform_in(dim, t, level, None, levels, filters, path=path)
                        want = a.copy()
                        ob.transform_in(want, level, dim, levels, filters)
                        assert_same(t.cpu().numpy(), want,
                                    f"{path} {extents} dim={dim} {levels} {len(filters)}p")


def test_reference_multipass_kat_on_gpu(pw):
    cases = json.load(open(os.path.join(GOLDEN, "plan_multipass.json")))
    for c in cases:
        a = np.array([float.fromhex(h) for h in c["input"]])
        filters = [(o, tuple(f)) for o, f in c["filters"]]
        for path in ("auto", "lds_line", "per_pass"):
            t = torch.from_numpy(a).cuda()
            pw.transform_in(0, t, [c["level"]], [c["maximum_level"]], c["levels"],
                            filters, path=path)
            expected = np.array([float.fromhex(h) for h in c["expected"]])
            assert_same(t.cpu().numpy(), expected, f"{path} {c['passes']}:{c['mask']}")


def test_golden_regression_vectors_on_gpu(pw):
    manifest = json.load(open(os.path.join(GOLDEN, "regression.json")))
    for m in manifest:
        dt = np.dtype(m["dtype"])
        load = lambda tag: np.fromfile(
            os.path.join(GOLDEN, f"regression_{m['name']}_{tag}.raw"), dtype=dt
        ).reshape(m["extents"])
        a = load("input")
        fwd = gpu_named(pw, a, m["level"], m["minimum_level"], m["wavelet"], True)
        assert_same(fwd, load("hierarchized"), m["name"])
        back = gpu_named(pw, fwd, m["level"], m["minimum_level"], m["wavelet"], False)
        assert_same(back, load("dehierarchized"), m["name"])


def test_long_lines_lds_and_per_pass(pw):
    rng = np.random.default_rng(15)
    for dtype, L, extents_fn in ((np.float64, 12, lambda n: [3, n + 1]),
                                 (np.float64, 14, lambda n: [n]),
                                 (np.float64, 15, lambda n: [n + 1]),
                                 (np.float32, 13, lambda n: [n + 1, 3]),
                                 (np.float32, 16, lambda n: [n])):
        n = 1 << L
        extents = extents_fn(n)
        level = [L if e in (n, n + 1) else 1 for e in extents]
        for name in ("hat", "biorthogonal"):
            a = closed_random(rng, extents, level, dtype)
            lmin = [0] * len(level)
            got = gpu_named(pw, a, level, lmin, name, True)
            assert_same(got, oracle_named(a, level, lmin, name, True), f"L={L} {extents}")


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_long_ladder_every_period(pw, dtype):
    """The long-line kernel (lines beyond the fused kernel: float64 2^12..2^15,
    float32 2^13..2^16 points) on strided axes (3 and 33 columns, several
    tiles of lines with a partial last one) and the contiguous axis, with and
    without images, every wavelet, both directions, three minimum levels;
    the explicit path equals the LDS-line path where that one holds the line."""
    rng = np.random.default_rng(18)
    periods = range(12, 16) if dtype == np.float64 else range(13, 17)
    for L in periods:
        n = 1 << L
        for extents, level, dim in (([n + 1, 3], [L, 1], 0), ([n, 33], [L, 5], 0),
                                    ([3, n + 1], [1, L], 1), ([2, n], [1, L], 1)):
            for name in WAVELETS:
                lmin_L = {12: 0, 13: 6, 14: 13, 15: 0, 16: 9}[L]
                lmin = [lmin_L if d == dim else 0 for d in range(2)]
                a = closed_random(rng, extents, level, dtype)
                for fwd in (True, False):
                    got = gpu_named(pw, a, level, lmin, name, fwd)
                    want = oracle_named(a, level, lmin, name, fwd)
                    assert_same(got, want, f"L={L} {extents} {name} lmin={lmin} fwd={fwd}")
        # the explicit schedule, and the LDS-line kernel on the same line
        level, extents = [L, 2], [n + 1, 5]
        levels = list(range(L, 0, -1))
        a = closed_random(rng, extents, level, dtype)
        out = {}
        for path in ("long_ladder", "lds_line") if L <= (14 if dtype == np.float64 else 15) \
                else ("long_ladder",):
            t = torch.from_numpy(a).cuda()
            pw.transform_in(0, t, level, level, levels, ob.FORWARD["biorthogonal"], path=path)
            out[path] = t.cpu().numpy()
        want = a.copy()
        ob.transform_in(want, level, 0, levels, ob.FORWARD["biorthogonal"])
        for path, got in out.items():
            assert_same(got, want, f"{path} L={L}")


@pytest.mark.parametrize("dtype,L", [(np.float64, 12), (np.float32, 13)])
def test_long_ladder_wide_tiles(pw, dtype, L):
    """More than 2048 strided lines of 2^12 (float64) / 2^13 (float32) points
    take 1024-thread tiles (the tests above, with few lines, take 256-thread
    ones): both directions against the oracle."""
    rng = np.random.default_rng(20)
    n = 1 << L
    level, extents = [L, 11], [n + 1, 2049]
    for name in ("hat", "biorthogonal"):
        a = closed_random(rng, extents, level, dtype)
        lmin = [0, 11]
        fwd = gpu_named(pw, a, level, lmin, name, True)
        assert_same(fwd, oracle_named(a, level, lmin, name, True), f"{extents} {name} fwd")
        back = gpu_named(pw, fwd, level, lmin, name, False)
        assert_same(back, oracle_named(fwd, level, lmin, name, False), f"{extents} {name} inv")


@pytest.mark.parametrize("dtype,L", [(np.float64, 12), (np.float64, 13), (np.float32, 13),
                                     (np.float32, 14)])
def test_long_ladder_contiguous_staged(pw, dtype, L):
    """Contiguous lines the long-line kernel stages through LDS (float64 up
    to 2^13, float32 up to 2^14 points), a few lines per round: 2049 lines
    (1024-thread tiles, several rounds per tile, a partial last tile) and 5
    lines (256-thread tiles), with and without images, both directions."""
    rng = np.random.default_rng(21)
    n = 1 << L
    for lines, image in ((2049, True), (2049, False), (5, True)):
        extents = [lines, n + 1 if image else n]
        level = [11 if lines == 2049 else 2, L]
        lmin = [level[0], 0]
        for name in ("hat", "biorthogonal"):
            a = closed_random(rng, extents, level, dtype)
            fwd = gpu_named(pw, a, level, lmin, name, True)
            assert_same(fwd, oracle_named(a, level, lmin, name, True),
                        f"{extents} {name} fwd")
            back = gpu_named(pw, fwd, level, lmin, name, False)
            assert_same(back, oracle_named(fwd, level, lmin, name, False),
                        f"{extents} {name} inv")


@pytest.mark.parametrize("name", WAVELETS)
def test_long_ladder_column_tiles(pw, name):
    """Strided float64 lines of 256 points through the long-line kernel's
    column tiles (16-point chunks): 1024-thread tiles of 64 columns over
    16385 lines (the last tile holds one), over 16384 lines without images,
    256-thread tiles over 25 lines and over an inner axis (1161 lines, 129
    between consecutive points); the explicit path at two minimum levels in
    both directions, and the automatic schedule over every axis (it takes the
    column tiles for the two-pass tables on the 16384-line grids), against
    the oracle."""
    rng = np.random.default_rng(22)
    cases = (([8, 14], [257, 16385], 0), ([8, 14], [256, 16384], 0),
             ([8, 2, 2], [257, 5, 5], 0), ([3, 8, 7], [9, 257, 129], 1))
    for level, extents, dim in cases:
        a = closed_random(rng, extents, level, np.float64)
        for lmin_L in (0, 3):
            for fwd in (True, False):
                levels = list(range(8, lmin_L, -1)) if fwd else list(range(lmin_L + 1, 9))
                filters = (ob.FORWARD if fwd else ob.RECONSTRUCT)[name]
                t = torch.from_numpy(a).cuda()
                pw.transform_in(dim, t, level, level, levels, filters, path="long_ladder")
                want = a.copy()
                ob.transform_in(want, level, dim, levels, filters)
                assert_same(t.cpu().numpy(), want,
                            f"{extents} axis {dim} {name} lmin={lmin_L} fwd={fwd}")
        lmin = [0] * len(level)
        for fwd in (True, False):
            assert_same(gpu_named(pw, a, level, lmin, name, fwd),
                        oracle_named(a, level, lmin, name, fwd), f"{extents} {name} auto")


def test_auto_path_names_the_schedule_that_runs(pw):
    """The schedule paliwa_auto_path reports for an axis, requested
    explicitly, gives the automatic sweep's result bit for bit (every
    schedule does; this checks the report names a schedule that holds the
    line), on the 257^3 grid's axes for the three tables and on a grid whose
    strided lines go to the column tiles under the hat table."""
    rng = np.random.default_rng(23)
    for extents, level in (([257, 257, 257], [8, 8, 8]), ([257, 129, 1025], [8, 7, 10])):
        a = closed_random(rng, extents, level, np.float64)
        for name in WAVELETS:
            for dim in range(3):
                for fwd in (True, False):
                    path = pw.auto_path(dim, extents, level, [0, 0, 0], name, fwd)
                    levels = list(range(level[dim], 0, -1)) if fwd else list(range(1, level[dim] + 1))
                    filters = (ob.FORWARD if fwd else ob.RECONSTRUCT)[name]
                    t = torch.from_numpy(a).cuda()
                    (pw.hierarchize_in if fwd else pw.dehierarchize_in)(dim, t, level, [0, 0, 0],
                                                                       level, name)
                    u = torch.from_numpy(a).cuda()
                    pw.transform_in(dim, u, level, level, levels, filters, path=path)
                    assert torch.equal(t, u), (extents, name, dim, fwd, path)


def test_long_ladder_rejects_generic_lists(pw):
    """The long-line kernel runs ladders only: an explicit request with a
    non-consecutive level list is refused; the automatic schedule takes the
    LDS-line kernel for it, bit-exact."""
    rng = np.random.default_rng(19)
    level, extents = [12, 2], [4097, 5]
    a = closed_random(rng, extents, level, np.float64)
    t = torch.from_numpy(a).cuda()
    with pytest.raises(pw.UnsupportedError):
        pw.transform_in(0, t, level, level, [12, 10, 11], ob.FORWARD["hat"], path="long_ladder")
    pw.transform_in(0, t, level, level, [12, 10, 11], ob.FORWARD["hat"])
    want = a.copy()
    ob.transform_in(want, level, 0, [12, 10, 11], ob.FORWARD["hat"])
    assert_same(t.cpu().numpy(), want, "generic list on a long line")


def test_axes_without_levels(pw):
    """Axes whose minimum level equals their level sweep nothing: a grid
    with every axis so, or of level 0 only, comes back unchanged, and mixed
    grids match the oracle (which transforms only the other axes)."""
    rng = np.random.default_rng(17)
    cases = [
        ([5, 4], [33, 16], [5, 4]),          # no axis sweeps
        ([0, 0, 0], [2, 1, 2], [0, 0, 0]),   # level 0 everywhere
        ([5, 4, 3], [33, 16, 9], [5, 0, 3]),  # only the middle axis sweeps
        ([6, 3], [64, 9], [2, 3]),           # only the first axis sweeps
    ]
    for name in WAVELETS:
        for dtype in (np.float64, np.float32):
            for level, extents, lmin in cases:
                a = closed_random(rng, extents, level, dtype)
                for fwd in (True, False):
                    got = gpu_named(pw, a, level, lmin, name, fwd)
                    want = oracle_named(a, level, lmin, name, fwd)
                    assert_same(got, want, f"{extents} lmin={lmin} {name} fwd={fwd}")
                    if lmin == level:
                        assert np.array_equal(got, a)


def test_python_api_errors_on_gpu(pw):
    t = torch.zeros(17, 16, dtype=torch.float64, device="cuda")
    with pytest.raises(KeyError):
        pw.hierarchize(t, [4, 4], [0, 0], [4, 4], "haar")
    with pytest.raises(RuntimeError, match="Filter offset not supported"):
        pw.transform_in(0, t, [4, 4], [4, 4], [4], [(3, (0.0, 1.0, 0.0))])
    with pytest.raises(ValueError):
        pw.hierarchize(t, [4, 4], [5, 0], [4, 4])
    with pytest.raises(ValueError):
        pw.hierarchize(torch.zeros(18, 16, dtype=torch.float64, device="cuda"),
                       [4, 4], [0, 0], [4, 4])
    with pytest.raises(ValueError):
        pw.hierarchize(t.t(), [4, 4], [0, 0], [4, 4])  # not contiguous
