"""ctypes binding of the C ABI declared in include/paliwa_hip.h.

The library is built in-tree (paliwa_amd/lib/libpaliwa_hip.so, by
``__graft_entry__.build()`` or ``make -C paliwa_amd/csrc``).  Importing this
module loads it; a missing library raises immediately -- there is no fallback
implementation.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY_PATH = os.path.join(_HERE, "lib", "libpaliwa_hip.so")

PALIWA_OK = 0
PALIWA_ERR_INVALID_ARGUMENT = -1
PALIWA_ERR_UNKNOWN_WAVELET = -2
PALIWA_ERR_FILTER_OFFSET = -3
PALIWA_ERR_LEVEL = -4
PALIWA_ERR_SHAPE = -5
PALIWA_ERR_DTYPE = -6
PALIWA_ERR_UNSUPPORTED = -7
PALIWA_ERR_HIP = -8
PALIWA_ERR_IO = -9

PALIWA_FLOAT64 = 0
PALIWA_FLOAT32 = 1

PALIWA_MAX_DIMS = 26
PALIWA_MAX_FILTERS = 16
PALIWA_MAX_LEVELS = 64

PATHS = {"auto": 0, "fused": 1, "lds_line": 2, "per_pass": 3, "long_ladder": 4}

PALIWA_TUNE_FORCE = 1
PALIWA_TUNE_MEASURE = 2
PALIWA_TUNE_KEEP_BOXES = 4


class paliwa_filter(ctypes.Structure):
    _fields_ = [
        ("offset", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("coefficients", ctypes.c_double * 3),
    ]


class paliwa_periodic_bounds(ctypes.Structure):
    _fields_ = [("min", ctypes.c_int64), ("max", ctypes.c_int64)]


class paliwa_axis_storage(ctypes.Structure):
    _fields_ = [("count", ctypes.c_int64), ("front", ctypes.c_int64),
                ("step", ctypes.c_int64), ("coordinates", ctypes.c_void_p)]


_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_pbp = ctypes.POINTER(paliwa_periodic_bounds)
_fp = ctypes.POINTER(paliwa_filter)
_asp = ctypes.POINTER(paliwa_axis_storage)
_vp = ctypes.c_void_p
_i32 = ctypes.c_int32

# name -> (restype, argtypes); mirrors include/paliwa_hip.h
SIGNATURES = {
    "paliwa_last_error": (ctypes.c_char_p, []),
    "paliwa_version": (ctypes.c_char_p, []),
    "paliwa_wavelet_filters": (ctypes.c_int, [ctypes.c_char_p, _i32, _fp, _i32, _i32p]),
    "paliwa_transform_in": (
        ctypes.c_int,
        [_vp, _i32, _i32, _i64p, _i32p, _i32p, _i32, _i32p, _i32, _fp, _i32, _vp],
    ),
    "paliwa_transform_in_path": (
        ctypes.c_int,
        [_vp, _i32, _i32, _i64p, _i32p, _i32p, _i32, _i32p, _i32, _fp, _i32, _i32, _vp],
    ),
    "paliwa_auto_path": (
        ctypes.c_int,
        [_i32, _i32, _i64p, _i32p, _i32p, _i32, ctypes.c_char_p, _i32, _i32p],
    ),
    "paliwa_hierarchize_in": (
        ctypes.c_int,
        [_vp, _i32, _i32, _i64p, _i32p, _i32p, _i32p, _i32, ctypes.c_char_p, _vp],
    ),
    "paliwa_dehierarchize_in": (
        ctypes.c_int,
        [_vp, _i32, _i32, _i64p, _i32p, _i32p, _i32p, _i32, ctypes.c_char_p, _vp],
    ),
    "paliwa_hierarchize": (
        ctypes.c_int,
        [_vp, _i32, _i32, _i64p, _i32p, _i32p, _i32p, ctypes.c_char_p, _vp],
    ),
    "paliwa_dehierarchize": (
        ctypes.c_int,
        [_vp, _i32, _i32, _i64p, _i32p, _i32p, _i32p, ctypes.c_char_p, _vp],
    ),
    "paliwa_transform_in_checked": (
        ctypes.c_int,
        [_vp, _i32, _i32, _asp, _i32p, _i32p, _i32, _i32p, _i32, _fp, _i32, _vp],
    ),
    "paliwa_hierarchize_in_checked": (
        ctypes.c_int,
        [_vp, _i32, _i32, _asp, _i32p, _i32p, _i32p, _i32, ctypes.c_char_p, _vp],
    ),
    "paliwa_dehierarchize_in_checked": (
        ctypes.c_int,
        [_vp, _i32, _i32, _asp, _i32p, _i32p, _i32p, _i32, ctypes.c_char_p, _vp],
    ),
    "paliwa_hierarchize_checked": (
        ctypes.c_int,
        [_vp, _i32, _i32, _asp, _i32p, _i32p, _i32p, ctypes.c_char_p, _vp],
    ),
    "paliwa_dehierarchize_checked": (
        ctypes.c_int,
        [_vp, _i32, _i32, _asp, _i32p, _i32p, _i32p, ctypes.c_char_p, _vp],
    ),
    "paliwa_batch_create": (
        ctypes.c_int,
        [ctypes.POINTER(_vp), _i32, _i32, _i32, ctypes.POINTER(_vp), _i64p, _i32p,
         _i32p, _i32p],
    ),
    "paliwa_batch_hierarchize": (ctypes.c_int, [_vp, ctypes.c_char_p, _vp]),
    "paliwa_batch_dehierarchize": (ctypes.c_int, [_vp, ctypes.c_char_p, _vp]),
    "paliwa_batch_set_graphs": (ctypes.c_int, [_vp, _i32]),
    "paliwa_batch_info": (ctypes.c_int, [_vp, _i32p, _i64p, _i64p]),
    "paliwa_batch_set_boxes": (ctypes.c_int, [_vp, _i32, _i32, _i32]),
    "paliwa_batch_traffic": (ctypes.c_int, [_vp, _i64p]),
    "paliwa_batch_get_boxes": (ctypes.c_int, [_vp, _i32p, _i32p, _i32p]),
    "paliwa_batch_destroy": (ctypes.c_int, [_vp]),
    "paliwa_batch_tune": (ctypes.c_int, [_vp, ctypes.c_char_p, _i32, ctypes.c_char_p, _i32, _vp]),
    "paliwa_batch_tune_info": (
        ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                       _i32p, _i32p]),
    "paliwa_transform_bounds": (
        ctypes.c_int, [paliwa_periodic_bounds, _i32, _i32, _i32p, _i32, _fp, _i32, _pbp, _pbp]),
    "paliwa_covering_periodic_bounds": (ctypes.c_int, [_i64p, ctypes.c_int64, ctypes.c_int64,
                                                       _pbp]),
    "paliwa_partial_create": (
        ctypes.c_int, [ctypes.POINTER(_vp), _i32, _i32, _i64p, _i32, _i32, _i32p, _i32, _fp, _i32,
                       ctypes.c_int64, ctypes.c_int64, ctypes.c_int64]),
    "paliwa_partial_ghosts": (ctypes.c_int, [_vp, _i64p, _i64p]),
    "paliwa_partial_transform": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "paliwa_partial_destroy": (ctypes.c_int, [_vp]),
    "paliwa_dump_to_binary_file": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_char_p, _vp]),
    "paliwa_load_from_binary_file": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_char_p, _vp]),
    "paliwa_combine_create": (
        ctypes.c_int, [ctypes.POINTER(_vp), _i32, _i32, _i32, ctypes.POINTER(_vp), _i64p,
                       _i32p, ctypes.POINTER(ctypes.c_double), _i32p, _i32]),
    "paliwa_combine_info": (ctypes.c_int, [_vp, _i32p, _i64p, _i64p, _i64p]),
    "paliwa_combine_subspaces": (ctypes.c_int, [_vp, _i32p, _i64p, _i32p]),
    "paliwa_combine_gather": (ctypes.c_int, [_vp, _vp, _vp]),
    "paliwa_combine_scatter": (ctypes.c_int, [_vp, _vp, _vp, _i64p, _vp]),
    "paliwa_combine_destroy": (ctypes.c_int, [_vp]),
    "paliwa_learned_param_count": (ctypes.c_int, []),
    "paliwa_quantize_f32": (ctypes.c_int, [_vp, _vp, ctypes.c_int64, ctypes.c_float, _vp]),
    "paliwa_dequantize_f32": (ctypes.c_int, [_vp, _vp, ctypes.c_int64, ctypes.c_float, _vp]),
    "paliwa_learned_hierarchize": (
        ctypes.c_int, [_vp, _i32, _i64p, _i32p, _i32p, _i32, _vp, _vp, _i32, _i32, _vp, _vp,
                       _vp]),
    "paliwa_learned_dehierarchize": (
        ctypes.c_int, [_vp, _i32, _i64p, _i32p, _i32p, _i32, _vp, _vp, _i32, _i32, _vp, _vp]),
    "paliwa_learned_extract_pairs": (
        ctypes.c_int, [_vp, _i32, _i64p, _i32p, _i32p, _i32, _i32, _vp, _vp, ctypes.c_uint32,
                       ctypes.c_uint32, ctypes.c_uint32, _vp]),
    "paliwa_learned_mlp_eval": (ctypes.c_int, [_vp, _vp, _vp, _i32, _vp]),
    "paliwa_learned_gradient_workspace": (ctypes.c_int64, [_i32]),
    "paliwa_learned_gradient": (ctypes.c_int, [_vp, _i32, _vp, _i32, _vp, _vp, _vp, _vp, _vp]),
    "paliwa_learned_gradient_static": (ctypes.c_int, [_vp, _i32, _vp, _i32, _vp, _vp, _vp, _vp]),
    "paliwa_learned_adam": (
        ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                       ctypes.c_float, _i32, _vp]),
    "paliwa_learned_fill": (ctypes.c_int, [_vp, _i32, ctypes.c_float, _vp]),
}


def _load() -> ctypes.CDLL:
    if not os.path.exists(LIBRARY_PATH):
        raise ImportError(
            f"{LIBRARY_PATH} is missing: build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` or "
            "`make -C paliwa_amd/csrc` (there is no non-HIP fallback)"
        )
    lib = ctypes.CDLL(LIBRARY_PATH)
    for name, (restype, argtypes) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    return lib


lib = _load()


def i32_array(values):
    values = [int(v) for v in values]
    return (ctypes.c_int32 * max(1, len(values)))(*values)


def i64_array(values):
    values = [int(v) for v in values]
    return (ctypes.c_int64 * max(1, len(values)))(*values)


def filter_array(filters):
    arr = (paliwa_filter * max(1, len(filters)))()
    for i, (offset, coefficients) in enumerate(filters):
        arr[i].offset = int(offset)
        arr[i].reserved = 0
        for k in range(3):
            arr[i].coefficients[k] = float(coefficients[k])
    return arr


def last_error() -> str:
    msg = lib.paliwa_last_error()
    return msg.decode() if msg else ""
