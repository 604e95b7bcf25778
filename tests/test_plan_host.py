"""The partial-output planner (paliwa_transform_bounds, host code) against
the brute force of the stencil graph in the oracle, as the reference's
planner tests do (test/test_paliwa_plan.cpp:29-95, 199-221)."""
import ctypes

import numpy as np
import pytest

import oracle_bridge as ob
from paliwa_amd import _lib, partial

_check = ob.lib.oracle_plan_check
_check.restype = ctypes.c_int
_check.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
                   ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                   ctypes.c_int, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]


def check_plan(level, maximum, start, count, levels, offsets):
    period, spacing = 1 << maximum, 1 << (maximum - level)
    outputs = (start * spacing, ((start + count - 1) * spacing) % period) if count else (-1, -1)
    nl, nf = len(levels), len(offsets)
    filters = _lib.filter_array([(o, (0.0, 1.0, 0.0)) for o in offsets])
    writes = (_lib.paliwa_periodic_bounds * max(1, nl * nf))()
    required = (_lib.paliwa_periodic_bounds * (level + 1))()
    rc = _lib.lib.paliwa_transform_bounds(
        _lib.paliwa_periodic_bounds(*outputs), level, maximum, _lib.i32_array(levels), nl,
        filters, nf, writes, required)
    assert rc == 0, _lib.last_error()
    w = (ctypes.c_int64 * max(2, 2 * nl * nf))(*[v for b in writes[:nl * nf] for v in (b.min, b.max)])
    r = (ctypes.c_int64 * (2 * level + 2))(*[v for b in required for v in (b.min, b.max)])
    return _check(level, maximum, start, count, (ctypes.c_int * max(1, nl))(*levels), nl,
                  (ctypes.c_int * max(1, nf))(*offsets), nf, w, r)


def test_exact_access_sets():
    """Every parity sequence of up to five passes, every ordinary and
    wrapped output interval, both directions, partial levels, and a
    coarser input grid (test/test_paliwa_plan.cpp:70-95)."""
    failures = []
    for maximum in (4, 6):
        for passes in range(1, 6):
            for mask in range(1 << passes):
                offsets = [(mask >> p) & 1 for p in range(passes)]
                for levels in ([4, 3, 2, 1], [1, 2, 3, 4], [4, 3], [3, 4], [4], []):
                    for start in range(16):
                        for count in range(17):
                            rc = check_plan(4, maximum, start, count, levels, offsets)
                            if rc != 0:
                                failures.append((maximum, offsets, levels, start, count, rc))
    assert not failures, failures[:5]


def test_longer_filter_sequences():
    """Random level ranges and 6-12 pass parity sequences
    (test/test_paliwa_plan.cpp:97-112; a seeded numpy generator)."""
    rng = np.random.default_rng(20260928)
    for trial in range(2000):
        level = 1 + int(rng.integers(8))
        minimum = int(rng.integers(level + 1))
        levels = list(range(minimum + 1, level + 1))
        if rng.integers(2):
            levels.reverse()
        offsets = [int(x) for x in rng.integers(0, 2, size=6 + int(rng.integers(7)))]
        start = int(rng.integers(1 << level))
        count = int(rng.integers((1 << level) + 1))
        assert check_plan(level, level + 2, start, count, levels, offsets) == 0, trial


def test_nonconsecutive_levels_are_rejected():
    for levels in ([4, 1], [1, 4], [1, 2, 1]):
        with pytest.raises(Exception) as e:
            partial.transform_bounds((0, 0), 4, 4, levels, [(1, (0.0, 1.0, 0.0))])
        assert "consecutive" in str(e.value)
    with pytest.raises(Exception, match="Invalid transform level"):
        partial.transform_bounds((0, 0), 4, 4, [5, 4], [(1, (0.0, 1.0, 0.0))])


def test_periodic_pole_cover():
    """test/test_paliwa_plan.cpp:210-221."""
    cases = [([], (-1, -1)), ([12], (12, 12)), ([8, 12, 16], (8, 16)),
             ([0, 4, 56, 60], (56, 4)), ([0, 16, 32, 48], (0, 48))]
    for indices, want in cases:
        assert partial.covering_periodic_bounds(indices, 64) == want


def test_planner_output_form():
    """A hat pass at the finest level writes exactly the odd outputs and reads
    their even neighbours, which are outputs already: the required sets are
    the outputs split by hierarchical level.  A full-period request plans
    every lattice."""
    w, r = partial.transform_bounds((4, 12), 4, 4, [4], [(1, (-0.5, 1.0, -0.5))])
    assert w == [[(5, 11)]]
    assert r == [(-1, -1), (8, 8), (4, 12), (6, 10), (5, 11)]
    w, r = partial.transform_bounds((0, 15), 4, 4, [4, 3, 2, 1], [(1, (0.0, 1.0, 0.0))])
    assert w[0] == [(1, 15)] and w[3] == [(8, 8)]


def test_error_precedence():
    """Walking backwards, the reference meets the last level's range check
    before any pass's parity check (paliwa/paliwa_domains.hpp:407-414)."""
    bad = [(1, (0.0, 1.0, 0.0)), (2, (0.0, 1.0, 0.0))]
    with pytest.raises(Exception, match="Filter offset not supported"):
        partial.transform_bounds((0, 0), 4, 4, [5, 4], bad)
    with pytest.raises(Exception, match="Invalid transform level"):
        partial.transform_bounds((0, 0), 4, 4, [4, 5], bad)
    with pytest.raises(Exception, match="consecutive"):
        partial.transform_bounds((0, 0), 4, 4, [4, 2], bad)
    # no levels: no pass runs, so no parity is inspected
    w, r = partial.transform_bounds((3, 5), 4, 4, [], bad)
    assert w == [] and r[0] == (-1, -1)


def test_cover_ties_leave_out_the_wrapping_gap():
    assert partial.covering_periodic_bounds([0, 32], 64) == (0, 32)
    # two equal interior gaps: the first one is left out
    assert partial.covering_periodic_bounds([0, 10, 20, 21], 30) == (10, 0)
    assert partial.covering_periodic_bounds([5, 6], 8) == (5, 6)


def test_largest_period_saturates():
    """At the largest period (2^62) a deep multi-pass ladder makes every
    level's dependency cone cover its whole lattice; the runs must not
    overflow on the way.  Closed form: a whole lattice of level h >= 1 is
    {2^(62-h), 2^62 - 2^(62-h)}, level 0 is {0, 0}."""
    offsets = [1, 0] * 8
    levels = list(range(62, 0, -1))
    filters = [(o, (0.0, 1.0, 0.0)) for o in offsets]
    w, r = partial.transform_bounds((0, 0), 62, 62, levels, filters)
    assert r[0] == (0, 0)
    for h in range(1, 63):
        assert r[h] == (1 << (62 - h), (1 << 62) - (1 << (62 - h))), h
    # the coarsest level's passes write its whole lattice
    assert w[-1][0] == (1 << 61, 1 << 61) and w[-1][1] == (0, 0)
