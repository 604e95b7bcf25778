"""Python mirror of the reference's transform interface
(paliwa/paliwa_transform.hpp), over the HIP library's C ABI.

Same names, argument meaning, order and defaults as the C++ templates:

    hierarchize(grid, level, minimum_level, maximum_level, wavelet_name="hat")
        -- paliwa/paliwa_transform.hpp:401-415
    dehierarchize(grid, level, minimum_level, maximum_level, wavelet_name="hat")
        -- paliwa/paliwa_transform.hpp:417-431
    hierarchize_in(dim, grid, ...) / dehierarchize_in(dim, grid, ...)
        -- paliwa/paliwa_transform.hpp:346-399; the template parameter Dim is
           the first argument here
    transform_in(dim, grid, level, maximum_level, levels, filters)
        -- paliwa/paliwa_transform.hpp:215-251

Partial buffers (the reference's checked branch, taken for every buffer that
is not the complete strided lattice, paliwa/paliwa_transform.hpp:24-51,
224-248) go through the same names with a ``storage`` argument: one entry
per axis giving the finest-level coordinate of every position, either
``Affine(front, step)`` (DDC's DiscreteDomain / StridedDiscreteDomain) or a
1-D int64 tensor of ascending coordinates on the grid's device (DDC's
SparseDiscreteDomain):

    hierarchize_checked(grid, storage, level, minimum_level, maximum_level,
                        wavelet_name="hat")
    dehierarchize_checked(...), hierarchize_in_checked(dim, ...),
    dehierarchize_in_checked(dim, ...),
    transform_in_checked(dim, grid, storage, level, maximum_level, levels,
                         filters)

A grid is a contiguous float64/float32 torch tensor in GPU memory, row-major
with the last axis contiguous, of extent 2^level[d] (or 2^level[d]+1 with a
periodic image at the end) per axis.  The calls are asynchronous on the
tensor's device, on the current torch stream unless ``stream`` (a
torch.cuda.Stream or a raw hipStream_t integer) is given.

Errors raise the Python counterpart of the reference's exception:
std::out_of_range (unknown wavelet, ``map::at``) -> UnknownWaveletError
(a KeyError); std::runtime_error ("Filter offset not supported") ->
FilterOffsetError (a RuntimeError); the reference's level asserts ->
LevelError (a ValueError).
"""
from __future__ import annotations

import ctypes
from typing import Iterable, Sequence

import torch

from . import _lib

__all__ = [
    "PaliwaError",
    "UnknownWaveletError",
    "FilterOffsetError",
    "LevelError",
    "ShapeError",
    "UnsupportedError",
    "HipError",
    "hierarchize",
    "dehierarchize",
    "hierarchize_in",
    "dehierarchize_in",
    "transform_in",
    "Affine",
    "hierarchize_checked",
    "dehierarchize_checked",
    "hierarchize_in_checked",
    "dehierarchize_in_checked",
    "transform_in_checked",
    "wavelet_filters",
    "lifting_wavelet_filter_offsets_and_coefficients",
    "lifting_wavelet_reconstruct_offsets_and_coefficients",
]


class PaliwaError(Exception):
    """Base class of errors reported by the HIP library."""


class UnknownWaveletError(PaliwaError, KeyError):
    """Unknown wavelet name (std::out_of_range from std::map::at)."""


class FilterOffsetError(PaliwaError, RuntimeError):
    """Filter offset other than 0 or 1 (std::runtime_error)."""


class LevelError(PaliwaError, ValueError):
    """Level ranges violated (asserted by the reference)."""


class ShapeError(PaliwaError, ValueError):
    """Extent is neither 2^level nor 2^level+1, or wrong rank/layout."""


class UnsupportedError(PaliwaError, ValueError):
    """Beyond the library's size limits."""


class HipError(PaliwaError, RuntimeError):
    """A HIP runtime call failed."""


_ERRORS = {
    _lib.PALIWA_ERR_INVALID_ARGUMENT: ValueError,
    _lib.PALIWA_ERR_UNKNOWN_WAVELET: UnknownWaveletError,
    _lib.PALIWA_ERR_FILTER_OFFSET: FilterOffsetError,
    _lib.PALIWA_ERR_LEVEL: LevelError,
    _lib.PALIWA_ERR_SHAPE: ShapeError,
    _lib.PALIWA_ERR_DTYPE: ShapeError,
    _lib.PALIWA_ERR_UNSUPPORTED: UnsupportedError,
    _lib.PALIWA_ERR_HIP: HipError,
    _lib.PALIWA_ERR_IO: OSError,
}


def check(rc: int) -> None:
    """Raise the exception matching a C-ABI status code."""
    if rc == _lib.PALIWA_OK:
        return
    exc = _ERRORS.get(rc, PaliwaError)
    message = _lib.last_error()
    if exc is ValueError:
        raise ValueError(message)
    raise exc(message)


def wavelet_filters(wavelet_name: str, reconstruct: bool = False):
    """Filter list of a named wavelet: [(offset, (c0, c1, c2)), ...]
    (paliwa/paliwa_wavelets.hpp:15-28)."""
    out = (_lib.paliwa_filter * _lib.PALIWA_MAX_FILTERS)()
    count = ctypes.c_int32(0)
    check(
        _lib.lib.paliwa_wavelet_filters(
            wavelet_name.encode(), int(bool(reconstruct)), out,
            _lib.PALIWA_MAX_FILTERS, ctypes.byref(count)
        )
    )
    return [
        (out[i].offset, tuple(out[i].coefficients[k] for k in range(3)))
        for i in range(count.value)
    ]


WAVELET_NAMES = ("biorthogonal", "fullweighting", "hat")

lifting_wavelet_filter_offsets_and_coefficients = {
    name: wavelet_filters(name, False) for name in WAVELET_NAMES
}
lifting_wavelet_reconstruct_offsets_and_coefficients = {
    name: wavelet_filters(name, True) for name in WAVELET_NAMES
}


def _dtype_code(t: torch.Tensor) -> int:
    if t.dtype == torch.float64:
        return _lib.PALIWA_FLOAT64
    if t.dtype == torch.float32:
        return _lib.PALIWA_FLOAT32
    raise ShapeError(f"grid dtype must be float64 or float32, got {t.dtype}")


def _check_grid(grid: torch.Tensor, level: Sequence[int]) -> None:
    if not isinstance(grid, torch.Tensor):
        raise TypeError("grid must be a torch.Tensor in GPU memory")
    if grid.device.type != "cuda":
        raise ShapeError("grid must live in GPU memory (device 'cuda')")
    if not grid.is_contiguous():
        raise ShapeError("grid must be contiguous (row-major, last axis fastest)")
    if grid.dim() != len(level):
        raise ShapeError(
            f"grid has {grid.dim()} axes but level has {len(level)} entries"
        )


def _stream_handle(grid: torch.Tensor, stream) -> ctypes.c_void_p:
    if stream is None:
        stream = torch.cuda.current_stream(grid.device)
    if isinstance(stream, torch.cuda.Stream):
        return ctypes.c_void_p(stream.cuda_stream)
    return ctypes.c_void_p(int(stream))


def _named(fn, grid, level, minimum_level, maximum_level, wavelet_name,
           stream, dim=None):
    level = list(level)
    _check_grid(grid, level)
    args = [
        ctypes.c_void_p(grid.data_ptr()), _dtype_code(grid), grid.dim(),
        _lib.i64_array(grid.shape), _lib.i32_array(level),
        _lib.i32_array(minimum_level), _lib.i32_array(maximum_level),
    ]
    if dim is not None:
        args.append(int(dim))
    args += [str(wavelet_name).encode(), _stream_handle(grid, stream)]
    with torch.cuda.device(grid.device):
        check(fn(*args))


def hierarchize(grid, level, minimum_level, maximum_level,
                wavelet_name: str = "hat", stream=None) -> None:
    """Forward transform along axes 0..D-1, in place
    (paliwa/paliwa_transform.hpp:401-415)."""
    _named(_lib.lib.paliwa_hierarchize, grid, level, minimum_level,
           maximum_level, wavelet_name, stream)


def dehierarchize(grid, level, minimum_level, maximum_level,
                  wavelet_name: str = "hat", stream=None) -> None:
    """Inverse transform along axes 0..D-1, in place
    (paliwa/paliwa_transform.hpp:417-431)."""
    _named(_lib.lib.paliwa_dehierarchize, grid, level, minimum_level,
           maximum_level, wavelet_name, stream)


def hierarchize_in(dim: int, grid, level, minimum_level, maximum_level,
                   wavelet_name: str = "hat", stream=None) -> bool:
    """Forward transform along one axis (paliwa/paliwa_transform.hpp:346-372)."""
    _named(_lib.lib.paliwa_hierarchize_in, grid, level, minimum_level,
           maximum_level, wavelet_name, stream, dim=dim)
    return True


def dehierarchize_in(dim: int, grid, level, minimum_level, maximum_level,
                     wavelet_name: str = "hat", stream=None) -> bool:
    """Inverse transform along one axis (paliwa/paliwa_transform.hpp:374-399)."""
    _named(_lib.lib.paliwa_dehierarchize_in, grid, level, minimum_level,
           maximum_level, wavelet_name, stream, dim=dim)
    return True


def auto_path(dim: int, extents, level, minimum_level, wavelet_name: str = "hat",
              forward: bool = True, dtype=torch.float64) -> str:
    """The kernel schedule the automatic path takes for hierarchize_in
    (forward) or dehierarchize_in along ``dim`` of a grid of these extents:
    "fused", "long_ladder", "lds_line" or "per_pass", or "auto" when the axis
    has no level to sweep.  Launches nothing and needs no device (an
    introspection of this library's rules, with no counterpart in the
    reference)."""
    level = list(level)
    code = {torch.float64: _lib.PALIWA_FLOAT64, torch.float32: _lib.PALIWA_FLOAT32}.get(dtype)
    if code is None:
        raise TypeError("dtype must be torch.float64 or torch.float32")
    out = ctypes.c_int32(-1)
    check(_lib.lib.paliwa_auto_path(code, len(level), _lib.i64_array(extents),
                                    _lib.i32_array(level), _lib.i32_array(minimum_level), int(dim),
                                    wavelet_name.encode(), int(bool(forward)), ctypes.byref(out)))
    return {v: k for k, v in _lib.PATHS.items()}[out.value]


def transform_in(dim: int, grid, level, maximum_level, levels: Iterable[int],
                 filters, stream=None, path: str = "auto") -> bool:
    """Any level list and filter list along one axis
    (paliwa/paliwa_transform.hpp:215-251).  ``filters`` is a sequence of
    (offset, (c0, c1, c2)); ``path`` selects the kernel schedule ("auto",
    "fused", "lds_line", "per_pass")."""
    level = list(level)
    _check_grid(grid, level)
    levels = [int(l) for l in levels]
    filters = list(filters)
    with torch.cuda.device(grid.device):
        check(
            _lib.lib.paliwa_transform_in_path(
                ctypes.c_void_p(grid.data_ptr()), _dtype_code(grid),
                grid.dim(), _lib.i64_array(grid.shape), _lib.i32_array(level),
                None if maximum_level is None else _lib.i32_array(maximum_level),
                int(dim), _lib.i32_array(levels), len(levels),
                _lib.filter_array(filters), len(filters), _lib.PATHS[path],
                _stream_handle(grid, stream),
            )
        )
    return True


class Affine:
    """Coordinates front + p * step along one axis of a partial buffer."""

    def __init__(self, front: int, step: int = 1):
        self.front = int(front)
        self.step = int(step)


def _storage_array(grid: torch.Tensor, storage):
    storage = list(storage)
    if len(storage) != grid.dim():
        raise ShapeError(f"grid has {grid.dim()} axes but storage has {len(storage)}")
    arr = (_lib.paliwa_axis_storage * len(storage))()
    for d, s in enumerate(storage):
        arr[d].count = int(grid.shape[d])
        if isinstance(s, Affine):
            arr[d].front, arr[d].step, arr[d].coordinates = s.front, s.step, None
        else:
            if not (isinstance(s, torch.Tensor) and s.dtype == torch.int64 and s.dim() == 1
                    and s.is_contiguous() and s.device == grid.device):
                raise ShapeError("a coordinate list must be a contiguous 1-D int64 tensor "
                                 "on the grid's device")
            if s.numel() != grid.shape[d]:
                raise ShapeError(f"axis {d}: {s.numel()} coordinates for extent {grid.shape[d]}")
            arr[d].front, arr[d].step = 0, 1
            arr[d].coordinates = ctypes.c_void_p(s.data_ptr()) if s.numel() else None
    return arr


def _named_checked(fn, grid, storage, level, minimum_level, maximum_level,
                   wavelet_name, stream, dim=None):
    level = list(level)
    _check_grid(grid, level)
    args = [
        ctypes.c_void_p(grid.data_ptr()), _dtype_code(grid), grid.dim(),
        _storage_array(grid, storage), _lib.i32_array(level),
        _lib.i32_array(minimum_level), _lib.i32_array(maximum_level),
    ]
    if dim is not None:
        args.append(int(dim))
    args += [str(wavelet_name).encode(), _stream_handle(grid, stream)]
    with torch.cuda.device(grid.device):
        check(fn(*args))


def hierarchize_checked(grid, storage, level, minimum_level, maximum_level,
                        wavelet_name: str = "hat", stream=None) -> None:
    """hierarchize on a partial buffer (paliwa/paliwa_transform.hpp:401-415,
    checked branch :24-51)."""
    _named_checked(_lib.lib.paliwa_hierarchize_checked, grid, storage, level,
                   minimum_level, maximum_level, wavelet_name, stream)


def dehierarchize_checked(grid, storage, level, minimum_level, maximum_level,
                          wavelet_name: str = "hat", stream=None) -> None:
    """dehierarchize on a partial buffer (paliwa/paliwa_transform.hpp:417-431)."""
    _named_checked(_lib.lib.paliwa_dehierarchize_checked, grid, storage, level,
                   minimum_level, maximum_level, wavelet_name, stream)


def hierarchize_in_checked(dim: int, grid, storage, level, minimum_level, maximum_level,
                           wavelet_name: str = "hat", stream=None) -> bool:
    _named_checked(_lib.lib.paliwa_hierarchize_in_checked, grid, storage, level,
                   minimum_level, maximum_level, wavelet_name, stream, dim=dim)
    return True


def dehierarchize_in_checked(dim: int, grid, storage, level, minimum_level, maximum_level,
                             wavelet_name: str = "hat", stream=None) -> bool:
    _named_checked(_lib.lib.paliwa_dehierarchize_in_checked, grid, storage, level,
                   minimum_level, maximum_level, wavelet_name, stream, dim=dim)
    return True


def transform_in_checked(dim: int, grid, storage, level, maximum_level,
                         levels: Iterable[int], filters, stream=None) -> bool:
    """transform_in on a partial buffer (paliwa/paliwa_transform.hpp:215-251,
    checked branch)."""
    level = list(level)
    _check_grid(grid, level)
    levels = [int(l) for l in levels]
    filters = list(filters)
    with torch.cuda.device(grid.device):
        check(
            _lib.lib.paliwa_transform_in_checked(
                ctypes.c_void_p(grid.data_ptr()), _dtype_code(grid), grid.dim(),
                _storage_array(grid, storage), _lib.i32_array(level),
                _lib.i32_array(maximum_level), int(dim), _lib.i32_array(levels),
                len(levels), _lib.filter_array(filters), len(filters),
                _stream_handle(grid, stream),
            )
        )
    return True
