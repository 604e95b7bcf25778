"""Batched transform of a set of grids -- the combination technique's
component grids -- over the C ABI's paliwa_batch plan.

Mirrors the reference's per-grid loop
(test/test_paliwa_combination_technique.cpp:164-174)

    for grid, level in zip(grids, levels):
        paliwa.hierarchize(grid, level, minimum_level, maximum_level, wavelet)

with one object: ``CombinationBatch(grids, levels, minimum_level,
maximum_level).hierarchize(wavelet)``.  Each grid's result is bit-identical to
hierarchize on that grid alone.
"""
from __future__ import annotations

import ctypes
import itertools
from typing import Sequence

import torch

from . import _lib
from .transform import _dtype_code, _stream_handle, check


def combination_levels(dim: int, level_sums: Sequence[int], minimum: int = 1):
    """All level vectors l of length `dim` with every l_i >= `minimum` and
    |l|_1 in `level_sums` (a combination-technique index set), ordered by
    level sum, then lexicographically."""
    out = []
    for total in level_sums:
        for l in itertools.product(range(minimum, total - minimum * (dim - 1) + 1), repeat=dim):
            if sum(l) == total:
                out.append(tuple(l))
    return out


class CombinationBatch:
    """A plan over many grids (torch tensors in GPU memory, same dtype and
    rank, dense row-major, extent 2^l or 2^l+1 per axis)."""

    def __init__(self, grids, levels, minimum_level, maximum_level):
        grids = list(grids)
        levels = [list(l) for l in levels]
        if not grids or len(grids) != len(levels):
            raise ValueError("need one level vector per grid")
        ndim = grids[0].dim()
        dtype = _dtype_code(grids[0])
        for g in grids:
            if g.device.type != "cuda" or not g.is_contiguous() or g.dim() != ndim:
                raise ValueError("grids must be contiguous GPU tensors of one rank")
            if g.dtype != grids[0].dtype:
                raise ValueError("grids must share one dtype")
        self.grids = grids  # keep the memory alive as long as the plan
        self.device = grids[0].device
        n = len(grids)
        ptrs = (ctypes.c_void_p * n)(*[g.data_ptr() for g in grids])
        ext = _lib.i64_array([e for g in grids for e in g.shape])
        lv = _lib.i32_array([x for l in levels for x in l])
        self._plan = ctypes.c_void_p(None)
        with torch.cuda.device(self.device):
            check(_lib.lib.paliwa_batch_create(
                ctypes.byref(self._plan), dtype, ndim, n, ptrs, ext, lv,
                _lib.i32_array(minimum_level), _lib.i32_array(maximum_level)))

    def _run(self, fn, wavelet_name, stream):
        with torch.cuda.device(self.device):
            check(fn(self._plan, str(wavelet_name).encode(),
                     _stream_handle(self.grids[0], stream)))

    def hierarchize(self, wavelet_name: str = "hat", stream=None) -> None:
        self._run(_lib.lib.paliwa_batch_hierarchize, wavelet_name, stream)

    def dehierarchize(self, wavelet_name: str = "hat", stream=None) -> None:
        self._run(_lib.lib.paliwa_batch_dehierarchize, wavelet_name, stream)

    def use_graphs(self, enable: bool) -> None:
        check(_lib.lib.paliwa_batch_set_graphs(self._plan, int(bool(enable))))

    def set_boxes(self, capacity_bytes: int, width: int = 32, threads: int = 256) -> None:
        """Box schedule: each grid's axes in boxes of consecutive axes swept
        in LDS (capacity 40960, 81920 or 163840 bytes; ``width`` inner points
        per tile, a power of two; 256 or 512 threads); capacity 0 runs the
        per-axis groups.  See paliwa_batch_set_boxes."""
        with torch.cuda.device(self.device):
            check(_lib.lib.paliwa_batch_set_boxes(self._plan, int(capacity_bytes), int(width),
                                                  int(threads)))

    def boxes(self):
        """(capacity_bytes, width, threads) of the box schedule; capacity 0
        when the plan runs its per-axis groups."""
        cap, width, threads = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_int32(0)
        check(_lib.lib.paliwa_batch_get_boxes(self._plan, ctypes.byref(cap), ctypes.byref(width),
                                              ctypes.byref(threads)))
        return cap.value, width.value, threads.value

    def info(self):
        launches = ctypes.c_int32(0)
        points = ctypes.c_int64(0)
        nbytes = ctypes.c_int64(0)
        check(_lib.lib.paliwa_batch_info(self._plan, ctypes.byref(launches),
                                         ctypes.byref(points), ctypes.byref(nbytes)))
        read = ctypes.c_int64(0)
        check(_lib.lib.paliwa_batch_traffic(self._plan, ctypes.byref(read)))
        return {"launches_per_direction": launches.value, "points": points.value,
                "descriptor_bytes": nbytes.value, "points_read_per_direction": read.value}

    def tune(self, wavelet_name: str = "hat", forward: bool = True, cache=None,
             force: bool = False, measure: bool = False, keep_boxes: bool = False,
             stream=None):
        """Time alternative box schedules, tile widths and stream schedules on
        the plan's own grids (restored afterwards) and keep the fastest; see
        paliwa_batch_tune.  ``cache``: path of a tuning cache file;
        ``keep_boxes``: leave the plan's box schedule as it is.  Returns
        {"rule_ms", "tuned_ms", "evaluations", "streams"}."""
        flags = (_lib.PALIWA_TUNE_FORCE if force else 0) | (
            _lib.PALIWA_TUNE_MEASURE if measure else 0) | (
            _lib.PALIWA_TUNE_KEEP_BOXES if keep_boxes else 0)
        with torch.cuda.device(self.device):
            check(_lib.lib.paliwa_batch_tune(
                self._plan, str(wavelet_name).encode(), int(bool(forward)),
                None if cache is None else str(cache).encode(), flags,
                _stream_handle(self.grids[0], stream)))
        rule, tuned = ctypes.c_double(0), ctypes.c_double(0)
        evals, streams = ctypes.c_int32(0), ctypes.c_int32(0)
        check(_lib.lib.paliwa_batch_tune_info(self._plan, ctypes.byref(rule), ctypes.byref(tuned),
                                              ctypes.byref(evals), ctypes.byref(streams)))
        return {"rule_ms": rule.value, "tuned_ms": tuned.value, "evaluations": evals.value,
                "streams": streams.value}

    def close(self) -> None:
        if getattr(self, "_plan", None) is not None and self._plan.value:
            _lib.lib.paliwa_batch_destroy(self._plan)
            self._plan = ctypes.c_void_p(None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
