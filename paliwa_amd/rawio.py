"""Raw dumps in the reference's on-disk format: a grid's elements, row-major,
headerless, native byte order (paliwa/paliwa_io.hpp:13-23), with the file
names its combination-technique test and postprocess.py use
(test/test_paliwa_combination_technique.cpp:150-162, 321-331;
paliwa/postprocess.py:26-66).

Writing goes through the C ABI (paliwa_dump_to_binary_file) from device
memory; the readers use numpy only, so a dump can be inspected anywhere.  A
fixture manifest (JSON) lists raw files with their shape and dtype.
"""
from __future__ import annotations

import ctypes
import json
import os

import numpy as np

from . import _lib
from .transform import check


def level_str(level) -> str:
    """'l0_l1_..._Dd', the level part of the reference's file names."""
    level = [int(l) for l in level]
    return "_".join(str(l) for l in level) + f"_{len(level)}d"


def component_grid_filename(level) -> str:
    return f"strided_grid_{level_str(level)}.raw"


def full_grid_filename(maximum_level) -> str:
    return f"full_grid_{level_str(maximum_level)}.raw"


def distributed_grid_filename(level, rank: int, prefix: str = "distributed_grid") -> str:
    return f"{prefix}_{level_str(level)}_rank{int(rank)}.raw"


def dump(grid, filename: str, stream=None) -> None:
    """Write a contiguous GPU tensor's elements to `filename`."""
    import torch

    from .transform import _stream_handle

    if not grid.is_contiguous():
        raise ValueError("grid must be contiguous")
    with torch.cuda.device(grid.device):
        check(_lib.lib.paliwa_dump_to_binary_file(
            ctypes.c_void_p(grid.data_ptr()), grid.numel() * grid.element_size(),
            os.fsencode(filename), _stream_handle(grid, stream)))


def load(filename: str, shape, dtype=None, device="cuda", stream=None):
    """A GPU tensor of `shape` read from a raw dump."""
    import torch

    from .transform import _stream_handle

    dtype = torch.float64 if dtype is None else dtype
    out = torch.empty([int(s) for s in shape], dtype=dtype, device=device)
    with torch.cuda.device(out.device):
        check(_lib.lib.paliwa_load_from_binary_file(
            ctypes.c_void_p(out.data_ptr()), out.numel() * out.element_size(),
            os.fsencode(filename), _stream_handle(out, stream)))
    return out


def read_raw(filename: str, shape, dtype=np.float64) -> np.ndarray:
    """numpy only: a raw dump as an array of `shape`."""
    data = np.fromfile(filename, dtype=dtype)
    return data.reshape(tuple(int(s) for s in shape))


def load_serial_grid(level, directory: str = ".", dtype=np.float64):
    """A component grid dumped as strided_grid_<level>.raw (2^l points per
    axis), or None when the file is absent."""
    path = os.path.join(directory, component_grid_filename(level))
    if not os.path.exists(path):
        return None
    return read_raw(path, [1 << int(l) for l in level], dtype)


def load_distributed_grid(level, n_ranks_per_dim, directory: str = ".",
                          prefix: str = "distributed_grid", dtype=np.float64):
    """Reassemble a grid from per-rank dumps: rank r holds the block at its
    row-major Cartesian coordinates, equal blocks of 2^l / n points per axis.
    None when a file is absent."""
    global_points = [1 << int(l) for l in level]
    n_ranks_per_dim = [int(n) for n in n_ranks_per_dim]
    local_points = [g // n for g, n in zip(global_points, n_ranks_per_dim)]
    full = np.zeros(global_points, dtype=dtype)
    for rank in range(int(np.prod(n_ranks_per_dim))):
        path = os.path.join(directory, distributed_grid_filename(level, rank, prefix))
        if not os.path.exists(path):
            return None
        block = read_raw(path, local_points, dtype)
        coords, r = [], rank
        for n in reversed(n_ranks_per_dim):
            coords.insert(0, r % n)
            r //= n
        full[tuple(slice(c * lp, (c + 1) * lp) for c, lp in zip(coords, local_points))] = block
    return full


def write_manifest(path: str, entries) -> None:
    """A fixture manifest: a JSON list of {"file", "shape", "dtype", ...}
    records (further keys, e.g. level or wavelet, are kept as given)."""
    records = []
    for e in entries:
        rec = dict(e)
        rec["shape"] = [int(s) for s in rec["shape"]]
        rec["dtype"] = str(np.dtype(rec.get("dtype", "float64")))
        records.append(rec)
    with open(path, "w") as f:
        json.dump(records, f, indent=1)


def read_manifest(path: str):
    """(record, array) for each entry of a manifest; files are resolved next
    to the manifest."""
    base = os.path.dirname(os.path.abspath(path))
    with open(path) as f:
        records = json.load(f)
    return [(r, read_raw(os.path.join(base, r["file"]), r["shape"], np.dtype(r["dtype"])))
            for r in records]
