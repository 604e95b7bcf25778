"""Partial-output transform: the reference's transform_in with an `outputs`
domain (paliwa/paliwa_transform.hpp:253-344), its dependency planner
transform_bounds and covering_periodic_bounds (paliwa/paliwa_domains.hpp:
315-487), over the C ABI.

    plan = PartialTransform(block.shape, dim, level, levels, filters,
                            local_start, output_start, output_count)
    ghosts = full_axis_values_at(plan.ghost_indices)   # the caller's exchange
    plan(block, ghosts)                                # outputs of `block`

The planner functions run on the host only.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .transform import _dtype_code, _stream_handle, check

EMPTY = (-1, -1)


def _bounds(b):
    return (int(b.min), int(b.max))


def transform_bounds(outputs, level: int, maximum_level: int, levels, filters):
    """(writes, required): writes[i][p] = periodic bounds (min, max) of the
    points pass p of levels[i] writes; required[h] = bounds of the inputs read
    on hierarchical level h (h = 0..level).  Finest-level coordinates on the
    period 2^maximum_level; (-1, -1) is empty; min > max wraps."""
    levels = [int(l) for l in levels]
    filters = list(filters)
    out = _lib.paliwa_periodic_bounds(*outputs)
    nw = max(1, len(levels) * len(filters))
    writes = (_lib.paliwa_periodic_bounds * nw)()
    required = (_lib.paliwa_periodic_bounds * (int(level) + 1))()
    check(_lib.lib.paliwa_transform_bounds(
        out, int(level), int(maximum_level), _lib.i32_array(levels), len(levels),
        _lib.filter_array(filters), len(filters), writes, required))
    w = [[_bounds(writes[i * len(filters) + p]) for p in range(len(filters))]
         for i in range(len(levels))]
    return w, [_bounds(required[h]) for h in range(int(level) + 1)]


def covering_periodic_bounds(indices, period: int):
    """Smallest periodic interval holding the sorted, distinct indices."""
    idx = [int(i) for i in indices]
    cover = _lib.paliwa_periodic_bounds()
    check(_lib.lib.paliwa_covering_periodic_bounds(_lib.i64_array(idx), len(idx), int(period),
                                                   ctypes.byref(cover)))
    return _bounds(cover)


class PartialTransform:
    """A partial-output transform plan along axis `dim` of a block (dense
    row-major, `shape`) holding the run [local_start, local_start +
    shape[dim]) mod 2^level of an axis stored at `level`; outputs are the run
    [output_start, output_start + output_count) mod 2^level inside it.
    Calling it writes the outputs of the block; ghost values (the required
    inputs outside the local run, at `ghost_indices`) come in a second array
    shaped like the block with shape[dim] = len(ghost_indices)."""

    def __init__(self, shape, dim: int, level: int, levels, filters, local_start: int,
                 output_start: int, output_count: int, dtype=torch.float64, device="cuda"):
        self.shape = [int(x) for x in shape]
        self.dim = int(dim)
        self.dtype = dtype
        self.device = torch.device(device)
        probe = torch.empty(0, dtype=dtype)
        levels = [int(l) for l in levels]
        filters = list(filters)
        self._plan = ctypes.c_void_p(None)
        with torch.cuda.device(self.device):
            check(_lib.lib.paliwa_partial_create(
                ctypes.byref(self._plan), _dtype_code(probe), len(self.shape),
                _lib.i64_array(self.shape), self.dim, int(level), _lib.i32_array(levels),
                len(levels), _lib.filter_array(filters), len(filters), int(local_start),
                int(output_start), int(output_count)))
        n = ctypes.c_int64(0)
        check(_lib.lib.paliwa_partial_ghosts(self._plan, ctypes.byref(n), None))
        idx = (ctypes.c_int64 * max(1, n.value))()
        check(_lib.lib.paliwa_partial_ghosts(self._plan, ctypes.byref(n), idx))
        self.ghost_indices = list(idx)[:n.value]

    def ghost_shape(self):
        s = list(self.shape)
        s[self.dim] = len(self.ghost_indices)
        return s

    def __call__(self, block: torch.Tensor, ghosts: torch.Tensor | None = None,
                 stream=None) -> torch.Tensor:
        if list(block.shape) != self.shape or block.dtype != self.dtype or not block.is_contiguous():
            raise ValueError("block does not match the plan")
        if self.ghost_indices:
            if ghosts is None or list(ghosts.shape) != self.ghost_shape() or \
                    ghosts.dtype != self.dtype or not ghosts.is_contiguous():
                raise ValueError(f"ghosts of shape {self.ghost_shape()} are required")
        with torch.cuda.device(block.device):
            check(_lib.lib.paliwa_partial_transform(
                self._plan, ctypes.c_void_p(block.data_ptr()),
                ctypes.c_void_p(ghosts.data_ptr() if ghosts is not None else 0),
                _stream_handle(block, stream)))
        return block

    def close(self) -> None:
        if getattr(self, "_plan", None) is not None and self._plan.value:
            _lib.lib.paliwa_partial_destroy(self._plan)
            self._plan = ctypes.c_void_p(None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
