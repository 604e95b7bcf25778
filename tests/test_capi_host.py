"""Host-side checks of the HIP library's C ABI (no GPU needed).

The library must load, export every function include/paliwa_hip.h declares,
return the reference's filter tables, and reject invalid arguments with the
status codes that the C++ header maps onto the reference's exception types --
all before any HIP call, so these run on a machine without a GPU.
"""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

from paliwa_amd import _lib  # noqa: E402

FAKE = ctypes.c_void_p(0x1000)  # never dereferenced: every call below fails validation


def declared_functions():
    names = []
    for fname in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if fname.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fname)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"\b(paliwa_\w+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_function():
    names = declared_functions()
    assert "paliwa_hierarchize" in names and "paliwa_transform_in" in names
    for name in names:
        assert hasattr(_lib.lib, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"


def test_version():
    assert b"gfx950" in _lib.lib.paliwa_version()


def test_filter_tables_match_reference():
    from paliwa_amd import (lifting_wavelet_filter_offsets_and_coefficients as fwd,
                            lifting_wavelet_reconstruct_offsets_and_coefficients as rec)

    # paliwa/paliwa_wavelets.hpp:15-28
    assert fwd["hat"] == [(1, (-0.5, 1.0, -0.5))]
    assert fwd["biorthogonal"] == [(1, (-0.5, 1.0, -0.5)), (0, (0.25, 1.0, 0.25))]
    assert fwd["fullweighting"] == [(0, (0.25, 0.5, 0.25)), (1, (-0.5, 1.0, -0.5))]
    assert rec["hat"] == [(1, (0.5, 1.0, 0.5))]
    assert rec["biorthogonal"] == [(0, (-0.25, 1.0, -0.25)), (1, (0.5, 1.0, 0.5))]
    assert rec["fullweighting"] == [(1, (0.5, 1.0, 0.5)), (0, (-0.5, 2.0, -0.5))]


def test_unknown_wavelet_name():
    from paliwa_amd import UnknownWaveletError, wavelet_filters

    with pytest.raises(UnknownWaveletError):
        wavelet_filters("db4")
    with pytest.raises(KeyError):  # std::out_of_range from std::map::at
        wavelet_filters("haar")


def call_named(fn, ndim=2, extents=(17, 16), level=(4, 4), lmin=(0, 0),
               lmax=(4, 4), name=b"hat", dtype=_lib.PALIWA_FLOAT64, data=FAKE):
    return fn(data, dtype, ndim, _lib.i64_array(extents), _lib.i32_array(level),
              _lib.i32_array(lmin), _lib.i32_array(lmax), name, None)


@pytest.mark.parametrize("fn_name", ["paliwa_hierarchize", "paliwa_dehierarchize"])
def test_named_validation(fn_name):
    fn = getattr(_lib.lib, fn_name)
    assert call_named(fn, name=b"haar") == _lib.PALIWA_ERR_UNKNOWN_WAVELET
    assert _lib.last_error() == "map::at"
    assert call_named(fn, lmin=(5, 0)) == _lib.PALIWA_ERR_LEVEL
    assert call_named(fn, lmin=(-1, 0)) == _lib.PALIWA_ERR_LEVEL
    assert call_named(fn, lmax=(3, 4)) == _lib.PALIWA_ERR_LEVEL
    assert call_named(fn, extents=(15, 16)) == _lib.PALIWA_ERR_SHAPE
    assert call_named(fn, extents=(18, 16)) == _lib.PALIWA_ERR_SHAPE
    assert call_named(fn, dtype=7) == _lib.PALIWA_ERR_DTYPE
    assert call_named(fn, ndim=0) == _lib.PALIWA_ERR_UNSUPPORTED
    assert call_named(fn, ndim=27) == _lib.PALIWA_ERR_UNSUPPORTED
    assert call_named(fn, data=ctypes.c_void_p(0)) == _lib.PALIWA_ERR_INVALID_ARGUMENT


def test_named_in_validation():
    fn = _lib.lib.paliwa_hierarchize_in
    args = lambda dim, name=b"hat": (
        FAKE, _lib.PALIWA_FLOAT64, 2, _lib.i64_array((16, 16)),
        _lib.i32_array((4, 4)), _lib.i32_array((0, 0)), _lib.i32_array((4, 4)),
        dim, name, None)
    assert fn(*args(2)) == _lib.PALIWA_ERR_INVALID_ARGUMENT
    assert fn(*args(-1)) == _lib.PALIWA_ERR_INVALID_ARGUMENT
    assert fn(*args(0, b"")) == _lib.PALIWA_ERR_UNKNOWN_WAVELET


def call_transform_in(levels=(4, 3), filters=((1, (-0.5, 1.0, -0.5)),), dim=0,
                      level=(4, 4), extents=(16, 16), path=0):
    return _lib.lib.paliwa_transform_in_path(
        FAKE, _lib.PALIWA_FLOAT64, 2, _lib.i64_array(extents),
        _lib.i32_array(level), None, dim, _lib.i32_array(levels), len(levels),
        _lib.filter_array(list(filters)), len(filters), path, None)


def test_transform_in_validation():
    # paliwa/paliwa_transform.hpp:239-240
    assert call_transform_in(filters=((2, (0.0, 1.0, 0.0)),)) == _lib.PALIWA_ERR_FILTER_OFFSET
    assert _lib.last_error() == "Filter offset not supported"
    assert call_transform_in(filters=((-1, (0.0, 1.0, 0.0)),)) == _lib.PALIWA_ERR_FILTER_OFFSET
    # paliwa/paliwa_transform.hpp:232 asserts l > 0
    assert call_transform_in(levels=(0,)) == _lib.PALIWA_ERR_LEVEL
    assert call_transform_in(levels=(5,)) == _lib.PALIWA_ERR_LEVEL
    assert call_transform_in(dim=2) == _lib.PALIWA_ERR_INVALID_ARGUMENT
    assert call_transform_in(path=9) == _lib.PALIWA_ERR_INVALID_ARGUMENT
    many = tuple((1, (0.0, 1.0, 0.0)) for _ in range(17))
    assert call_transform_in(filters=many) == _lib.PALIWA_ERR_UNSUPPORTED
    assert call_transform_in(levels=tuple([1] * 65)) == _lib.PALIWA_ERR_UNSUPPORTED


def test_python_exceptions_map_status_codes():
    import paliwa_amd as p
    from paliwa_amd.transform import check

    for code, exc in (
        (_lib.PALIWA_ERR_UNKNOWN_WAVELET, KeyError),
        (_lib.PALIWA_ERR_FILTER_OFFSET, RuntimeError),
        (_lib.PALIWA_ERR_LEVEL, ValueError),
        (_lib.PALIWA_ERR_SHAPE, ValueError),
        (_lib.PALIWA_ERR_HIP, RuntimeError),
    ):
        with pytest.raises(exc):
            check(code)
    with pytest.raises(p.FilterOffsetError):
        check(_lib.PALIWA_ERR_FILTER_OFFSET)


def test_python_api_rejects_host_tensors():
    import torch
    import paliwa_amd as p

    with pytest.raises(p.ShapeError):
        p.hierarchize(torch.zeros(16, 16, dtype=torch.float64), [4, 4], [0, 0], [4, 4])


def test_batch_create_validation():
    plan = ctypes.c_void_p(None)
    ptrs = (ctypes.c_void_p * 2)(0x1000, 0x2000)
    ext = _lib.i64_array([17, 16, 9, 8])
    lv = _lib.i32_array([4, 4, 3, 3])
    mn, mx = _lib.i32_array([0, 0]), _lib.i32_array([4, 4])
    create = _lib.lib.paliwa_batch_create
    bad_ext = _lib.i64_array([17, 16, 10, 8])
    assert create(ctypes.byref(plan), 0, 2, 2, ptrs, bad_ext, lv, mn, mx) == _lib.PALIWA_ERR_SHAPE
    assert "grid 1" in _lib.last_error()
    bad_lv = _lib.i32_array([4, 4, 3, 5])
    assert create(ctypes.byref(plan), 0, 2, 2, ptrs, ext, bad_lv, mn, mx) in (
        _lib.PALIWA_ERR_SHAPE, _lib.PALIWA_ERR_LEVEL)
    assert create(ctypes.byref(plan), 0, 2, 0, ptrs, ext, lv, mn, mx) == _lib.PALIWA_ERR_INVALID_ARGUMENT
    assert create(ctypes.byref(plan), 9, 2, 2, ptrs, ext, lv, mn, mx) == _lib.PALIWA_ERR_DTYPE
    assert plan.value is None
    assert _lib.lib.paliwa_batch_hierarchize(None, b"hat", None) == _lib.PALIWA_ERR_INVALID_ARGUMENT


def test_combine_create_validation():
    plan = ctypes.c_void_p(None)
    ptrs = (ctypes.c_void_p * 2)(0x1000, 0x2000)
    ext = _lib.i64_array([17, 16, 9, 8])
    lv = _lib.i32_array([4, 4, 3, 3])
    coefs = (ctypes.c_double * 2)(1.0, -1.0)
    mx = _lib.i32_array([4, 4])
    create = _lib.lib.paliwa_combine_create
    bad_ext = _lib.i64_array([17, 16, 10, 8])
    assert create(ctypes.byref(plan), 0, 2, 2, ptrs, bad_ext, lv, coefs, mx, 2) == _lib.PALIWA_ERR_SHAPE
    assert "grid 1" in _lib.last_error()
    low_max = _lib.i32_array([4, 3])
    assert create(ctypes.byref(plan), 0, 2, 2, ptrs, ext, lv, coefs, low_max, 2) == _lib.PALIWA_ERR_LEVEL
    assert create(ctypes.byref(plan), 0, 2, 2, ptrs, ext, lv, coefs, mx, 0) == _lib.PALIWA_ERR_INVALID_ARGUMENT
    assert create(ctypes.byref(plan), 0, 9, 2, ptrs, ext, lv, coefs, mx, 2) == _lib.PALIWA_ERR_UNSUPPORTED
    assert create(ctypes.byref(plan), 7, 2, 2, ptrs, ext, lv, coefs, mx, 2) == _lib.PALIWA_ERR_DTYPE
    assert create(ctypes.byref(plan), 0, 2, 2, ptrs, ext, lv, None, mx, 2) == _lib.PALIWA_ERR_INVALID_ARGUMENT
    assert plan.value is None
    assert _lib.lib.paliwa_combine_gather(None, None, None) == _lib.PALIWA_ERR_INVALID_ARGUMENT
    assert _lib.lib.paliwa_combine_scatter(None, None, None, None, None) == _lib.PALIWA_ERR_INVALID_ARGUMENT
    assert _lib.lib.paliwa_combine_destroy(None) == 0


def test_combination_levels():
    from paliwa_amd import combination_levels

    # test/test_paliwa_combination_technique.cpp: the 4-D set of config 3
    levels = combination_levels(4, (13, 14, 15, 16), 1)
    assert len(levels) == 220 + 286 + 364 + 455
    assert len(set(levels)) == len(levels)
    assert all(min(l) >= 1 and sum(l) in (13, 14, 15, 16) for l in levels)


def test_axis_level_above_30_is_rejected_before_any_launch():
    """Lines are indexed with 32-bit integers, so a stored extent of 2^31
    points or more (a 2 x (2^31 + 1) grid is about 32 GB) is refused on the
    host with PALIWA_ERR_UNSUPPORTED; level 30 still passes the level check
    (it fails only on the fake pointer's shape here)."""
    for lv in (31, 40):
        ext = (2, (1 << lv) + 1)
        assert call_named(_lib.lib.paliwa_hierarchize, extents=ext, level=(1, lv),
                          lmax=(1, lv)) == _lib.PALIWA_ERR_UNSUPPORTED
        assert "above 30" in _lib.last_error()
        assert call_transform_in(levels=(lv,), dim=1, level=(1, lv),
                                 extents=ext) == _lib.PALIWA_ERR_UNSUPPORTED
    plan = ctypes.c_void_p(None)
    ptrs = (ctypes.c_void_p * 1)(0x1000)
    create = _lib.lib.paliwa_batch_create
    assert create(ctypes.byref(plan), 0, 2, 1, ptrs, _lib.i64_array([2, (1 << 31) + 1]),
                  _lib.i32_array([1, 31]), _lib.i32_array([0, 0]),
                  _lib.i32_array([1, 31])) == _lib.PALIWA_ERR_UNSUPPORTED
    assert plan.value is None
    # level 30 is accepted by the level check; a wrong extent is reported
    assert call_named(_lib.lib.paliwa_hierarchize, extents=(2, 7), level=(1, 30),
                      lmax=(1, 30)) == _lib.PALIWA_ERR_SHAPE


def test_checked_storage_validation():
    """Partial-buffer entry points validate their storage on the host."""
    S = _lib.paliwa_axis_storage

    def storage(*axes):
        arr = (S * len(axes))()
        for i, (count, front, step, coords) in enumerate(axes):
            arr[i].count, arr[i].front, arr[i].step, arr[i].coordinates = count, front, step, coords
        return arr

    def call(st, level=(3, 3), lmax=(4, 4), lmin=(0, 0), name=b"hat", ndim=2, dtype=0):
        return _lib.lib.paliwa_hierarchize_checked(
            FAKE, dtype, ndim, st, _lib.i32_array(level), _lib.i32_array(lmin),
            _lib.i32_array(lmax), name, None)

    ok_axis = (8, 0, 2, None)
    assert call(storage(ok_axis, (16, 0, 0, None))) == _lib.PALIWA_ERR_INVALID_ARGUMENT  # step 0
    assert call(storage(ok_axis, (16, -1, 1, None))) == _lib.PALIWA_ERR_INVALID_ARGUMENT  # front < 0
    assert call(storage(ok_axis, (17, 0, 1, None))) == _lib.PALIWA_ERR_SHAPE  # past the period
    assert call(storage(ok_axis, (8, 10, 1, None))) == _lib.PALIWA_ERR_SHAPE
    assert call(storage(ok_axis, ((1 << 30) + 2, 0, 1, 0x1000))) == _lib.PALIWA_ERR_UNSUPPORTED
    assert call(storage(ok_axis, ok_axis), level=(3, 5)) == _lib.PALIWA_ERR_LEVEL
    assert call(storage(ok_axis, ok_axis), lmax=(4, 63)) == _lib.PALIWA_ERR_LEVEL
    assert call(storage(ok_axis, ok_axis), lmin=(4, 0)) == _lib.PALIWA_ERR_LEVEL
    assert call(storage(ok_axis, ok_axis), name=b"haar") == _lib.PALIWA_ERR_UNKNOWN_WAVELET
    assert call(storage(ok_axis, ok_axis), dtype=5) == _lib.PALIWA_ERR_DTYPE
    assert call(None) == _lib.PALIWA_ERR_INVALID_ARGUMENT
    # an empty buffer is valid and does nothing
    assert call(storage(ok_axis, (0, 0, 1, None))) == _lib.PALIWA_OK
    fn = _lib.lib.paliwa_transform_in_checked
    filt = _lib.filter_array([(2, (0.0, 1.0, 0.0))])
    st = storage(ok_axis, ok_axis)
    assert fn(FAKE, 0, 2, st, _lib.i32_array([3, 3]), _lib.i32_array([4, 4]), 0,
              _lib.i32_array([3]), 1, filt, 1, None) == _lib.PALIWA_ERR_FILTER_OFFSET
    good = _lib.filter_array([(1, (0.0, 1.0, 0.0))])
    assert fn(FAKE, 0, 2, st, _lib.i32_array([3, 3]), _lib.i32_array([4, 4]), 0,
              _lib.i32_array([4]), 1, good, 1, None) == _lib.PALIWA_ERR_LEVEL
    assert fn(FAKE, 0, 2, st, _lib.i32_array([3, 3]), _lib.i32_array([4, 4]), 2,
              _lib.i32_array([3]), 1, good, 1, None) == _lib.PALIWA_ERR_INVALID_ARGUMENT


def test_entry_points_open_roctx_ranges():
    """The library brackets its entry points in roctx ranges (named per axis
    for the named transforms), which rocprofv3 --marker-trace shows."""
    import subprocess

    out = subprocess.run(["nm", "-D", _lib.LIBRARY_PATH], capture_output=True, text=True,
                         check=True).stdout
    assert "roctxRangePushA" in out and "roctxRangePop" in out
    strings = subprocess.run(["strings", _lib.LIBRARY_PATH], capture_output=True, text=True,
                             check=True).stdout
    for name in ("paliwa hierarchize axis ", "paliwa batch hierarchize", "paliwa combine gather"):
        assert name in strings, name
