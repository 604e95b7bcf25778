"""The automatic schedule's choice of kernel per axis (paliwa_auto_path; host
only, nothing is launched): the rules measured on the GPU, pinned on the
shapes they were measured on."""
import pytest
import torch

import paliwa_amd as pw


def path(dim, extents, level, lmin=None, name="biorthogonal", forward=True, dtype=torch.float64):
    return pw.auto_path(dim, extents, level, lmin if lmin is not None else [0] * len(level),
                        name, forward, dtype)


@pytest.mark.parametrize("forward", [True, False])
def test_column_tiles_on_the_257_cube(forward):
    """Strided float64 lines of 256 points with a two-pass table take the
    long-line kernel's column tiles; the contiguous axis, the one-pass hat
    table (except on an outer axis of a grid beyond the cache) and float32
    stay on the fused kernel."""
    cube, lv = [257] * 3, [8] * 3
    for name in ("biorthogonal", "fullweighting"):
        assert path(0, cube, lv, name=name, forward=forward) == "long_ladder"
        assert path(1, cube, lv, name=name, forward=forward) == "long_ladder"
        assert path(2, cube, lv, name=name, forward=forward) == "fused"
    assert path(0, cube, lv, name="hat", forward=forward) == "fused"
    assert path(1, cube, lv, name="hat", forward=forward) == "fused"
    assert path(0, cube, lv, dtype=torch.float32, forward=forward) == "fused"
    # the hat table takes them only on an outer axis of a grid beyond the cache
    assert path(0, [257, 1025, 1025], [8, 10, 10], name="hat", forward=forward) == "long_ladder"


def test_column_tile_rule_bounds():
    # an outer axis (4096 or more lines between consecutive points) at any size
    assert path(0, [257, 1025, 1025], [8, 10, 10]) == "long_ladder"
    # an inner axis: only while the grid is at most 160 MiB
    assert path(1, [257, 257, 129], [8, 8, 7]) == "long_ladder"     # 68 MB
    assert path(1, [257, 257, 513], [8, 8, 9]) == "fused"           # 271 MB
    assert path(0, [257, 257, 513], [8, 8, 9]) == "long_ladder"
    # fewer than 16384 lines (no 1024-thread tiles)
    assert path(0, [257, 5, 5], [8, 2, 2]) == "fused"
    assert path(0, [256, 16384], [8, 14]) == "long_ladder"
    assert path(0, [256, 8192], [8, 13]) == "fused"
    # other periods stay with their kernels
    assert path(0, [513] * 3, [9] * 3) == "fused"
    assert path(0, [129, 129, 129, 17], [7, 7, 7, 4]) == "fused"
    assert path(0, [2049, 2049], [11, 11]) == "fused"


def test_long_lines():
    assert path(0, [4097, 4097], [12, 12]) == "long_ladder"       # strided
    assert path(1, [4097, 4097], [12, 12]) == "long_ladder"       # staged contiguous
    assert path(1, [2, 16385], [1, 14]) == "lds_line"             # contiguous, not staged
    assert path(1, [2, 32769], [1, 15]) == "long_ladder"          # beyond the LDS-line kernel
    assert path(0, [65537, 2], [16, 1]) == "per_pass"             # beyond the long-line kernel
    assert path(0, [65537, 2], [16, 1], dtype=torch.float32) == "long_ladder"


def test_nothing_to_sweep_and_errors():
    assert path(0, [257] * 3, [8] * 3, lmin=[8, 0, 0]) == "auto"
    with pytest.raises(ValueError, match="dim out of range"):
        path(3, [257] * 3, [8] * 3)
    with pytest.raises(pw.UnknownWaveletError):
        path(0, [257] * 3, [8] * 3, name="haar")
    with pytest.raises(pw.ShapeError):
        path(0, [255, 257, 257], [8] * 3)
    with pytest.raises(pw.LevelError):
        path(0, [257] * 3, [8] * 3, lmin=[9, 0, 0])
