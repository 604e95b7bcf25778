"""ctypes access to the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is the checker: tests compare the HIP library against it; nothing
in the product imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "build", "liboracle.so")

FORWARD = {
    "hat": [(1, (-0.5, 1.0, -0.5))],
    "biorthogonal": [(1, (-0.5, 1.0, -0.5)), (0, (0.25, 1.0, 0.25))],
    "fullweighting": [(0, (0.25, 0.5, 0.25)), (1, (-0.5, 1.0, -0.5))],
}
RECONSTRUCT = {
    "hat": [(1, (0.5, 1.0, 0.5))],
    "biorthogonal": [(0, (-0.25, 1.0, -0.25)), (1, (0.5, 1.0, 0.5))],
    "fullweighting": [(1, (0.5, 1.0, 0.5)), (0, (-0.5, 2.0, -0.5))],
}

SERIAL, OMP_DIRECT, OMP_PACKED_POLES = 0, 1, 2


def _load():
    if not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    lib = ctypes.CDLL(LIB_PATH)
    i32p = ctypes.POINTER(ctypes.c_int32)
    i64p = ctypes.POINTER(ctypes.c_int64)
    lib.oracle_transform_in.restype = ctypes.c_int
    lib.oracle_transform_in.argtypes = [
        ctypes.c_void_p, ctypes.c_int, ctypes.c_int, i64p, i32p, ctypes.c_int,
        i32p, ctypes.c_int, i32p, ctypes.POINTER(ctypes.c_double), ctypes.c_int,
        ctypes.c_int,
    ]
    lib.oracle_transform.restype = ctypes.c_int
    lib.oracle_transform.argtypes = [
        ctypes.c_void_p, ctypes.c_int, ctypes.c_int, i64p, i32p, i32p, i32p,
        ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
    ]
    lib.oracle_transform_in_checked.restype = ctypes.c_int
    lib.oracle_transform_in_checked.argtypes = [
        ctypes.c_void_p, ctypes.c_int, ctypes.c_int, i64p, i64p, i32p, i32p, ctypes.c_int,
        i32p, ctypes.c_int, i32p, ctypes.POINTER(ctypes.c_double), ctypes.c_int,
    ]
    lib.oracle_last_error.restype = ctypes.c_char_p
    lib.oracle_num_threads.restype = ctypes.c_int
    return lib


lib = _load()


def _i32(v):
    v = [int(x) for x in v]
    return (ctypes.c_int32 * max(1, len(v)))(*v)


def _i64(v):
    v = [int(x) for x in v]
    return (ctypes.c_int64 * max(1, len(v)))(*v)


def _dtype(a: np.ndarray) -> int:
    if a.dtype == np.float64:
        return 0
    if a.dtype == np.float32:
        return 1
    raise TypeError(a.dtype)


class OracleError(Exception):
    def __init__(self, code, message):
        super().__init__(f"oracle error {code}: {message}")
        self.code = code


def _check(rc):
    if rc != 0:
        raise OracleError(rc, lib.oracle_last_error().decode())


def transform(a: np.ndarray, level, lmin, lmax, name: str, forward: bool,
              schedule: int = SERIAL) -> np.ndarray:
    """hierarchize (forward) / dehierarchize in place on a C-contiguous array."""
    assert a.flags.c_contiguous
    _check(lib.oracle_transform(
        a.ctypes.data, _dtype(a), a.ndim, _i64(a.shape), _i32(level), _i32(lmin),
        _i32(lmax), name.encode(), int(forward), schedule))
    return a


def hierarchize(a, level, lmin, lmax, name="hat", schedule=SERIAL):
    return transform(a, level, lmin, lmax, name, True, schedule)


def dehierarchize(a, level, lmin, lmax, name="hat", schedule=SERIAL):
    return transform(a, level, lmin, lmax, name, False, schedule)


def transform_in(a: np.ndarray, level, dim: int, levels, filters,
                 schedule: int = SERIAL) -> np.ndarray:
    assert a.flags.c_contiguous
    offsets = [f[0] for f in filters]
    coefs = [c for f in filters for c in f[1]]
    carr = (ctypes.c_double * max(1, len(coefs)))(*coefs)
    _check(lib.oracle_transform_in(
        a.ctypes.data, _dtype(a), a.ndim, _i64(a.shape), _i32(level), int(dim),
        _i32(levels), len(levels), _i32(offsets), carr, len(filters), schedule))
    return a


def transform_in_checked(a: np.ndarray, coords, level, maximum_level, dim: int, levels,
                         filters) -> np.ndarray:
    """The checked branch on a partial buffer: coords[d] = the ascending
    finest-level coordinates of the positions of axis d."""
    assert a.flags.c_contiguous and a.ndim == len(coords)
    counts = [len(c) for c in coords]
    assert list(a.shape) == counts
    flat = [int(x) for c in coords for x in c]
    offsets = [f[0] for f in filters]
    coefs = [c for f in filters for c in f[1]]
    carr = (ctypes.c_double * max(1, len(coefs)))(*coefs)
    _check(lib.oracle_transform_in_checked(
        a.ctypes.data, _dtype(a), a.ndim, _i64(counts), _i64(flat), _i32(level),
        _i32(maximum_level), int(dim), _i32(levels), len(levels), _i32(offsets), carr,
        len(filters)))
    return a


def transform_checked(a, coords, level, lmin, lmax, name, forward):
    """hierarchize / dehierarchize of a partial buffer: axes 0..D-1, levels
    decreasing (forward) or increasing, the named table."""
    filters = (FORWARD if forward else RECONSTRUCT)[name]
    for d in range(a.ndim):
        levels = list(range(level[d], lmin[d], -1)) if forward else \
            list(range(lmin[d] + 1, level[d] + 1))
        transform_in_checked(a, coords, level, lmax, d, levels, filters)
    return a


def num_threads() -> int:
    return lib.oracle_num_threads()
