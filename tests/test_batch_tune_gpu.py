"""Autotuning of a batched plan (paliwa_batch_tune) on the GPU: the grids
are restored bit for bit, the tuned plan's results are bit-identical to the
rule-based plan's, and a cache file written by one plan answers every choice
of the next without timing."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pw():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    import paliwa_amd

    return paliwa_amd


def grids_for(levels, dtype, seed):
    rng = np.random.default_rng(seed)
    out = []
    for l in levels:
        a = rng.uniform(-1, 1, size=[(1 << x) + 1 for x in l]).astype(dtype)
        for d, x in enumerate(l):
            src = [slice(None)] * len(l)
            dst = [slice(None)] * len(l)
            src[d], dst[d] = 0, 1 << x
            a[tuple(dst)] = a[tuple(src)]
        out.append(np.ascontiguousarray(a))
    return out


@pytest.mark.parametrize("dtype", [
    np.float64,
    pytest.param(np.float32, marks=pytest.mark.skip(
        reason="tuning of float32 plans is switched off: a float32 tuning run ended in a "
               "host segmentation fault inside paliwa_batch_tune whose cause is not found")),
])
def test_tuned_plan_is_bit_identical(pw, dtype, tmp_path):
    # lines of 512 and 1024 points on strided axes: groups whose tile width the
    # tuner chooses, next to groups of short lines
    levels = [l for l in pw.combination_levels(3, (11, 12), 1) if max(l) <= 10]
    lmin, lmax = [0, 0, 0], [10, 10, 10]
    arrays = grids_for(levels, dtype, seed=5)
    cache = tmp_path / "tune.txt"
    for name in ("hat", "biorthogonal"):
        tuned = [torch.from_numpy(a).cuda() for a in arrays]
        plan = pw.CombinationBatch(tuned, levels, lmin, lmax)
        report = plan.tune(name, forward=True, cache=cache, force=True, measure=True)
        torch.cuda.synchronize()
        assert report["evaluations"] >= 3 and report["rule_ms"] > 0 and report["tuned_ms"] > 0
        for a, g in zip(arrays, tuned):  # the grids are restored
            assert np.array_equal(g.cpu().numpy(), a)
        plain = [torch.from_numpy(a).cuda() for a in arrays]
        rule = pw.CombinationBatch(plain, levels, lmin, lmax)
        plan.hierarchize(name)
        rule.hierarchize(name)
        plan.dehierarchize(name)
        rule.dehierarchize(name)
        plan.hierarchize(name)
        rule.hierarchize(name)
        torch.cuda.synchronize()
        for g, h in zip(tuned, plain):
            assert np.array_equal(g.cpu().numpy(), h.cpu().numpy())
        text = cache.read_text()
        assert f"{name} forward" in text and "boxes " in text
        # a second plan over the same shapes takes every choice from the cache
        again = pw.CombinationBatch([torch.from_numpy(a).cuda() for a in arrays], levels, lmin,
                                    lmax)
        second = again.tune(name, forward=True, cache=cache)
        assert second["evaluations"] == 0 and second["tuned_ms"] < 0
        assert second["streams"] == report["streams"]
        plan.close()
        rule.close()
        again.close()


def test_group_widths_tuned_under_kept_boxes(pw, tmp_path):
    """With the per-axis groups kept (no box schedule), the tuner chooses the
    strided groups' tile widths and the stream schedule: bit-identical
    results, and a cache that answers a second plan."""
    levels = [l for l in pw.combination_levels(3, (11, 12), 1) if max(l) <= 10]
    lmin, lmax = [0, 0, 0], [10, 10, 10]
    arrays = grids_for(levels, np.float64, seed=6)
    cache = tmp_path / "tune.txt"
    tuned = [torch.from_numpy(a).cuda() for a in arrays]
    plan = pw.CombinationBatch(tuned, levels, lmin, lmax)
    plan.set_boxes(0)
    report = plan.tune("biorthogonal", forward=False, cache=cache, force=True, measure=True,
                       keep_boxes=True)
    assert report["evaluations"] >= 3 and plan.boxes()[0] == 0
    keys = [line.split()[0] for line in cache.read_text().splitlines()
            if line and not line.startswith("#")]
    assert "width" in keys and "schedule" in keys and "boxes" not in keys
    plain = [torch.from_numpy(a).cuda() for a in arrays]
    rule = pw.CombinationBatch(plain, levels, lmin, lmax)
    for p in (plan, rule):
        p.hierarchize("biorthogonal")
        p.dehierarchize("biorthogonal")
    torch.cuda.synchronize()
    for g, h in zip(tuned, plain):
        assert np.array_equal(g.cpu().numpy(), h.cpu().numpy())
    again = pw.CombinationBatch([torch.from_numpy(a).cuda() for a in arrays], levels, lmin, lmax)
    again.set_boxes(0)
    second = again.tune("biorthogonal", forward=False, cache=cache, keep_boxes=True)
    assert second["evaluations"] == 0 and second["streams"] == report["streams"]
    for p in (plan, rule, again):
        p.close()


def test_float32_tuning_is_refused(pw):
    levels = [(2, 3), (3, 2)]
    grids = [torch.zeros([(1 << x) + 1 for x in l], dtype=torch.float32, device="cuda")
             for l in levels]
    plan = pw.CombinationBatch(grids, levels, [0, 0], [3, 3])
    with pytest.raises(pw.UnsupportedError, match="float32"):
        plan.tune("hat")
    plan.close()


def test_tune_rejects_bad_arguments(pw):
    from paliwa_amd import _lib

    assert _lib.lib.paliwa_batch_tune(None, b"hat", 1, None, 0, None) == \
        _lib.PALIWA_ERR_INVALID_ARGUMENT
    levels = [(2, 3), (3, 2)]
    grids = [torch.zeros([(1 << x) + 1 for x in l], dtype=torch.float64, device="cuda")
             for l in levels]
    plan = pw.CombinationBatch(grids, levels, [0, 0], [3, 3])
    with pytest.raises(pw.UnknownWaveletError):
        plan.tune("haar")
    plan.close()
