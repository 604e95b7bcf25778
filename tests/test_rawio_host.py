"""Raw dump file names and the numpy-only readers (paliwa_amd/rawio.py)
against the reference's conventions (test/test_paliwa_combination_technique
.cpp:150-162, 321-331; paliwa/postprocess.py:26-66)."""
import json
import os

import numpy as np

from paliwa_amd import rawio

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_file_names():
    assert rawio.level_str([2, 11]) == "2_11_2d"
    assert rawio.component_grid_filename([4, 6, 7]) == "strided_grid_4_6_7_3d.raw"
    assert rawio.full_grid_filename([10, 11]) == "full_grid_10_11_2d.raw"
    assert rawio.distributed_grid_filename([5, 5], 3) == "distributed_grid_5_5_2d_rank3.raw"


def test_serial_reader(tmp_path):
    level = [3, 2]
    a = np.arange(32, dtype=np.float64).reshape(8, 4) * 0.5
    a.tofile(tmp_path / rawio.component_grid_filename(level))
    got = rawio.load_serial_grid(level, str(tmp_path))
    assert np.array_equal(got, a)
    assert rawio.load_serial_grid([1, 1], str(tmp_path)) is None


def test_distributed_reassembly(tmp_path):
    """Rank r holds the block at its row-major Cartesian coordinates."""
    level, ranks = [3, 2], [2, 2]
    full = np.random.default_rng(0).standard_normal((8, 4))
    for r in range(4):
        ci, cj = divmod(r, 2)
        full[ci * 4:(ci + 1) * 4, cj * 2:(cj + 1) * 2].tofile(
            tmp_path / rawio.distributed_grid_filename(level, r))
    assert np.array_equal(rawio.load_distributed_grid(level, ranks, str(tmp_path)), full)
    os.remove(tmp_path / rawio.distributed_grid_filename(level, 3))
    assert rawio.load_distributed_grid(level, ranks, str(tmp_path)) is None


def test_manifest_round_trip(tmp_path):
    a = np.linspace(0, 1, 12, dtype=np.float32).reshape(3, 4)
    a.tofile(tmp_path / "a.raw")
    rawio.write_manifest(str(tmp_path / "m.json"),
                         [{"file": "a.raw", "shape": (3, 4), "dtype": "float32", "level": [1, 2]}])
    [(rec, arr)] = rawio.read_manifest(str(tmp_path / "m.json"))
    assert rec["level"] == [1, 2] and rec["dtype"] == "float32"
    assert np.array_equal(arr, a)


def test_golden_fixtures_are_raw_dumps():
    """The committed regression fixtures are in this format."""
    with open(os.path.join(GOLDEN, "regression.json")) as f:
        cases = json.load(f)
    for c in cases:
        path = os.path.join(GOLDEN, f"regression_{c['name']}_input.raw")
        arr = rawio.read_raw(path, c["extents"], np.dtype(c["dtype"]))
        assert arr.shape == tuple(c["extents"])
