"""Partial-output transform on the GPU: a rank's block plus ghost values
against the full transform of the oracle (bit-exact outputs, untouched
ghosts and non-output points), as the reference's remote-storage tests do
(test/test_paliwa_plan.cpp:114-197, 286-360)."""
import numpy as np
import pytest
import torch

import oracle_bridge as ob

pytestmark = pytest.mark.gpu

FIVE_PASS = [(1, (-0.5, 1.0, 0.125)), (0, (0.0, 1.0, 0.25)), (0, (0.125, 1.0, 0.125)),
             (1, (0.0, 1.0, 0.25)), (1, (-0.25, 1.0, 0.125))]


@pytest.fixture(scope="module")
def pp():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    from paliwa_amd import partial

    return partial


def _cases():
    for name in ("hat", "biorthogonal", "fullweighting"):
        yield name, True, ob.FORWARD[name]
        yield name, False, ob.RECONSTRUCT[name]
    yield "five-pass", True, FIVE_PASS
    yield "five-pass", False, FIVE_PASS


def _run(pp, full, level, dim, levels, filters, local_start, count, out_off, out_count,
         dtype):
    """Transform the block [local_start, local_start + count) of axis `dim`
    with outputs at block offsets [out_off, out_off + out_count)."""
    L = level[dim]
    P = 1 << L
    run = [(local_start + j) % P for j in range(count)]
    block = np.ascontiguousarray(np.take(full, run, axis=dim))
    plan = pp.PartialTransform(block.shape, dim, L, levels, filters, local_start,
                               (local_start + out_off) % P, out_count, dtype=dtype)
    ghosts = None
    if plan.ghost_indices:
        gh = np.ascontiguousarray(np.take(full, plan.ghost_indices, axis=dim))
        ghosts = torch.from_numpy(gh).cuda()
        assert not set(plan.ghost_indices) & set(run)
    t = torch.from_numpy(block.copy()).cuda()
    plan(t, ghosts)
    got = t.cpu().numpy()
    if ghosts is not None:
        assert np.array_equal(ghosts.cpu().numpy(), gh)
    plan.close()
    return block, got, run


@pytest.mark.parametrize("shape_levels,dim", [((9,), 0), ((7, 5), 0), ((7, 5), 1),
                                              ((4, 6, 3), 1), ((4, 6, 3), 2), ((3, 5), 0),
                                              ((9, 4), 0), ((10, 3), 0), ((11, 2), 0)])
def test_blocks_with_ghosts(pp, shape_levels, dim):
    level = list(shape_levels)
    L = level[dim]
    P = 1 << L
    rng = np.random.default_rng(L * 10 + dim)
    for name, forward, filters in _cases():
        for lmin in (0, L // 2):
            levels = list(range(L, lmin, -1)) if forward else list(range(lmin + 1, L + 1))
            if not levels:
                continue
            for dtype in (np.float64, np.float32):
                full = rng.standard_normal([1 << l for l in level]).astype(dtype)
                want = full.copy()
                ob.transform_in(want, level, dim, levels, filters)
                blocks = [(0, P // 2), (P // 2, P // 2), (P - 3, min(P, 5)), (1, max(1, P // 4)),
                          (0, P)]
                for start, count in blocks:
                    for out_off, out_count in ((0, count), (count // 3, max(1, count // 3))):
                        out_count = min(out_count, count - out_off)
                        block, got, run = _run(pp, full, level, dim, levels, filters, start,
                                               count, out_off, out_count,
                                               torch.float64 if dtype == np.float64 else torch.float32)
                        exp = np.take(want, run, axis=dim)
                        sl = [slice(None)] * len(level)
                        sl[dim] = slice(out_off, out_off + out_count)
                        sl = tuple(sl)
                        assert np.array_equal(got[sl], exp[sl]), (name, forward, lmin, start, count,
                                                                  out_off, out_count, dtype)
                        rest = np.ones(got.shape, dtype=bool)
                        rest[sl] = False
                        assert np.array_equal(got[rest], block[rest])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_many_tiles_per_workgroup(pp, dtype):
    """Slabs of a 256^3 grid along axis 0 (65536 lines, several tiles per
    resident workgroup, so the next tile's rows are fetched while the current
    one computes): the whole-period ladder (biorthogonal down to level 0, a
    window of the whole period) and a short window (hat), slabs with and
    without wrap."""
    level, dim, L, P = [8, 8, 8], 0, 8, 256
    rng = np.random.default_rng(41)
    full = rng.standard_normal([P, P, P]).astype(dtype)
    tdtype = torch.float64 if dtype == np.float64 else torch.float32
    for name, lmin in (("biorthogonal", 0), ("hat", 0)):
        filters = ob.FORWARD[name]
        levels = list(range(L, lmin, -1))
        want = full.copy()
        ob.transform_in(want, level, dim, levels, filters)
        for start, count, out_off, out_count in ((0, 32, 0, 32), (240, 40, 5, 30)):
            block, got, run = _run(pp, full, level, dim, levels, filters, start, count,
                                   out_off, out_count, tdtype)
            exp = np.take(want, run, axis=dim)
            assert np.array_equal(got[out_off:out_off + out_count],
                                  exp[out_off:out_off + out_count]), (name, start, dtype)
            assert np.array_equal(got[:out_off], block[:out_off])
            assert np.array_equal(got[out_off + out_count:], block[out_off + out_count:])


def test_single_allocation_form(pp):
    """The whole axis held locally: no ghosts; outputs anywhere, wrapping."""
    level, dim = [6, 4], 0
    L, P = 6, 64
    rng = np.random.default_rng(7)
    full = rng.standard_normal([64, 16])
    filters = ob.FORWARD["biorthogonal"]
    levels = list(range(L, 0, -1))
    want = full.copy()
    ob.transform_in(want, level, dim, levels, filters)
    for out_start, out_count in ((0, 64), (60, 8), (17, 1), (5, 30)):
        plan = pp.PartialTransform(full.shape, dim, L, levels, filters, 0, out_start, out_count)
        assert plan.ghost_indices == []
        t = torch.from_numpy(full.copy()).cuda()
        plan(t)
        got = t.cpu().numpy()
        idx = [(out_start + k) % P for k in range(out_count)]
        assert np.array_equal(got[idx], want[idx])
        rest = sorted(set(range(P)) - set(idx))
        assert np.array_equal(got[rest], full[rest])


def test_plan_errors(pp):
    import paliwa_amd as pw

    filters = ob.FORWARD["hat"]
    with pytest.raises(ValueError):  # outputs leave the local run [0, 4)
        pp.PartialTransform([4, 4], 0, 3, [3, 2, 1], filters, 0, 3, 3)
    with pytest.raises(ValueError, match="consecutive"):
        pp.PartialTransform([8, 4], 0, 3, [3, 1], filters, 0, 0, 8)
    with pytest.raises(pw.UnsupportedError):
        pp.PartialTransform([1 << 14, 2], 0, 14, list(range(14, 0, -1)), filters, 0, 0, 1 << 14)
