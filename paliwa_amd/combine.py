"""Combination-technique combine over the C ABI's paliwa_combine plan: the
steps after the batched hierarchize in the reference's workflow
(test/test_paliwa_combination_technique.cpp:176-313).

    batch = CombinationBatch(grids, levels, minimum_level, maximum_level)
    batch.hierarchize(wavelet)                           # :164-174
    comb = SparseGridCombination(grids, levels, coefficients, maximum_level)
    sparse = comb.gather()                               # :254-290
    full = comb.scatter(sparse)                          # :292-313
    dehierarchize(full, maximum_level, minimum_level, maximum_level, wavelet)

``combine_full`` runs the last three lines.  Subspaces kept: those contained
in at least ``min_count`` grids (the reference keeps count > 1: min_count=2).
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .transform import _dtype_code, _stream_handle, check, dehierarchize


class SparseGridCombination:
    """A combine plan over component grids (torch tensors in GPU memory, one
    dtype and rank, dense row-major, extent 2^l or 2^l+1 per axis)."""

    def __init__(self, grids, levels, coefficients, maximum_level, min_count: int = 2):
        grids = list(grids)
        levels = [list(l) for l in levels]
        if not grids or len(grids) != len(levels) or len(grids) != len(coefficients):
            raise ValueError("need one level vector and one coefficient per grid")
        ndim = grids[0].dim()
        for g in grids:
            if g.device.type != "cuda" or not g.is_contiguous() or g.dim() != ndim:
                raise ValueError("grids must be contiguous GPU tensors of one rank")
            if g.dtype != grids[0].dtype:
                raise ValueError("grids must share one dtype")
        self.grids = grids  # the plan reads them: keep them alive
        self.levels = levels
        self.maximum_level = list(maximum_level)
        self.ndim = ndim
        self.dtype = grids[0].dtype
        self.device = grids[0].device
        n = len(grids)
        ptrs = (ctypes.c_void_p * n)(*[g.data_ptr() for g in grids])
        coefs = (ctypes.c_double * n)(*[float(c) for c in coefficients])
        self._plan = ctypes.c_void_p(None)
        with torch.cuda.device(self.device):
            check(_lib.lib.paliwa_combine_create(
                ctypes.byref(self._plan), _dtype_code(grids[0]), ndim, n, ptrs,
                _lib.i64_array([e for g in grids for e in g.shape]),
                _lib.i32_array([x for l in levels for x in l]), coefs,
                _lib.i32_array(self.maximum_level), int(min_count)))
        info = self.info()
        self.sparse_points = info["sparse_points"]
        ns = info["subspaces"]
        lv = (ctypes.c_int32 * max(1, ns * ndim))()
        off = (ctypes.c_int64 * (ns + 1))()
        cnt = (ctypes.c_int32 * max(1, ns))()
        check(_lib.lib.paliwa_combine_subspaces(self._plan, lv, off, cnt))
        self.subspaces = [tuple(lv[i * ndim:(i + 1) * ndim]) for i in range(ns)]
        self.offsets = list(off)
        self.counts = list(cnt)[:ns]

    def info(self):
        ns = ctypes.c_int32(0)
        pts = ctypes.c_int64(0)
        terms = ctypes.c_int64(0)
        nbytes = ctypes.c_int64(0)
        check(_lib.lib.paliwa_combine_info(self._plan, ctypes.byref(ns), ctypes.byref(pts),
                                           ctypes.byref(terms), ctypes.byref(nbytes)))
        return {"subspaces": ns.value, "sparse_points": pts.value, "terms": terms.value,
                "table_bytes": nbytes.value}

    def gather(self, out: torch.Tensor | None = None, stream=None) -> torch.Tensor:
        """Sparse-grid storage of the kept subspaces (1-D, sparse_points)."""
        if out is None:
            out = torch.empty(self.sparse_points, dtype=self.dtype, device=self.device)
        if out.numel() != self.sparse_points or out.dtype != self.dtype or not out.is_contiguous():
            raise ValueError("out must be a contiguous tensor of sparse_points elements")
        with torch.cuda.device(self.device):
            check(_lib.lib.paliwa_combine_gather(self._plan, ctypes.c_void_p(out.data_ptr()),
                                                 _stream_handle(out, stream)))
        return out

    def scatter(self, sparse: torch.Tensor, full: torch.Tensor | None = None,
                image: bool = False, stream=None) -> torch.Tensor:
        """The full grid at the maximum level (extent 2^M, or 2^M+1 with
        `image`), zero outside the kept subspaces."""
        if sparse.numel() != self.sparse_points or sparse.dtype != self.dtype:
            raise ValueError("sparse does not belong to this plan")
        if full is None:
            shape = [(1 << m) + (1 if image else 0) for m in self.maximum_level]
            full = torch.empty(shape, dtype=self.dtype, device=self.device)
        if full.dtype != self.dtype or not full.is_contiguous() or full.dim() != self.ndim:
            raise ValueError("full must be a contiguous tensor of the plan's dtype and rank")
        with torch.cuda.device(self.device):
            check(_lib.lib.paliwa_combine_scatter(
                self._plan, ctypes.c_void_p(sparse.data_ptr()), ctypes.c_void_p(full.data_ptr()),
                _lib.i64_array(full.shape), _stream_handle(full, stream)))
        return full

    def combine_full(self, minimum_level, wavelet_name: str = "hat", image: bool = False,
                     stream=None) -> torch.Tensor:
        """gather, scatter, and dehierarchize the full grid at the maximum
        level: the combined solution in nodal values."""
        sparse = self.gather(stream=stream)
        full = self.scatter(sparse, image=image, stream=stream)
        dehierarchize(full, self.maximum_level, minimum_level, self.maximum_level,
                      wavelet_name, stream=stream)
        return full

    def close(self) -> None:
        if getattr(self, "_plan", None) is not None and self._plan.value:
            _lib.lib.paliwa_combine_destroy(self._plan)
            self._plan = ctypes.c_void_p(None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
