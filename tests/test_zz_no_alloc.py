"""paliwa_learned_gradient_static (no workspace argument; partial sums in the
library's static device array) against paliwa_learned_gradient with a
caller workspace, through the C ABI: identical gradients and losses (the
same kernels, the partial sums added in the same order), on two calls in a
row, the second smaller, so that stale partial sums of the first call would
show; and the row limit refused on the host."""
import ctypes

import numpy as np
import pytest
import torch

from paliwa_amd import _lib

pytestmark = pytest.mark.gpu

STRIDE = 10  # 8 inputs, the target, a key


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def gradients(lib, pairs, index, count, weights, static):
    n = lib.paliwa_learned_param_count()
    grad = torch.zeros(n, dtype=torch.float32, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    if static:
        rc = lib.paliwa_learned_gradient_static(ptr(pairs), STRIDE, ptr(index), count,
                                                ptr(weights), ptr(grad), ptr(loss), None)
    else:
        ws = torch.empty(lib.paliwa_learned_gradient_workspace(count), dtype=torch.float32,
                         device="cuda")
        rc = lib.paliwa_learned_gradient(ptr(pairs), STRIDE, ptr(index), count, ptr(weights),
                                         ptr(grad), ptr(loss), ptr(ws), None)
    assert rc == 0, _lib.last_error()
    torch.cuda.synchronize()
    return grad.cpu(), loss.cpu()


def test_static_gradient_matches_workspace_gradient():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    lib = _lib.lib
    rng = np.random.default_rng(77)
    rows = 16384 + 100
    pairs = torch.from_numpy(rng.uniform(-1, 1, size=(rows, STRIDE)).astype(np.float32)).cuda()
    weights = torch.from_numpy(
        rng.normal(0, 0.3, size=lib.paliwa_learned_param_count()).astype(np.float32)).cuda()
    perm = torch.from_numpy(rng.permutation(rows).astype(np.int32)).cuda()
    # the largest call, then smaller ones, with and without a row index
    for count, index in ((16384, None), (3001, perm[:3001]), (64, None), (16384, perm[:16384]),
                         (1, None)):
        print(f"learned gradient, {count} rows, index {index is not None}", flush=True)
        ga, la = gradients(lib, pairs, index, count, weights, static=False)
        gb, lb = gradients(lib, pairs, index, count, weights, static=True)
        assert torch.isfinite(ga).all() and float(la[0]) > 0
        assert torch.equal(ga, gb), count
        assert torch.equal(la, lb), count


def test_row_limit_is_refused_before_any_launch():
    lib = _lib.lib
    fake = ctypes.c_void_p(0x1000)
    assert lib.paliwa_learned_gradient_static(fake, STRIDE, None, 16385, fake, fake, fake,
                                              None) == _lib.PALIWA_ERR_UNSUPPORTED
    assert "16384" in _lib.last_error()
    assert lib.paliwa_learned_gradient_static(fake, 8, None, 10, fake, fake, fake,
                                              None) == _lib.PALIWA_ERR_INVALID_ARGUMENT
