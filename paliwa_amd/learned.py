"""Learned lifting: an integer lifting transform whose predict step at levels
>= 3 adds a small MLP correction to the interpolet (hat) prediction.  An
extension on top of the plain method, which stays the fallback (per axis and
level) and the reference point for detail sparsity.  The HIP kernels are declared in
include/paliwa_hip.h (paliwa_learned_*); this module drives them.

Pipeline:
    q, scale = quantize(field)                        # int32 fixed point
    pairs = extract_pairs(training_grids, level)      # from the interpolet transform
    weights, losses = train(pairs)                    # hand-written fwd/bwd/Adam kernels
    masks, _ = choose_masks(q, level, lmin, weights)  # MLP only where it lowers L1
    forward(q, ...); inverse(q, ...)                  # exact integer round trip
"""
from __future__ import annotations

import ctypes
import math
from typing import Sequence

import numpy as np
import torch

from . import _lib
from .transform import check

PARAMS = _lib.lib.paliwa_learned_param_count()
QUANT_BITS = 22  # |q| <= 2^22: headroom for detail growth over three axes
MAX_LEVEL = 40
FEATURES_RAW = 0        # inputs c_k / max|c|
FEATURES_DETRENDED = 1  # inputs: deviation from the interpolet line / its max-abs


def _ptr(t: torch.Tensor | None):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _stream(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def init_weights(seed: int = 0, device="cuda") -> torch.Tensor:
    """Packed fp32 parameters; the output layer starts at zero, so the initial
    predictor is the interpolet itself."""
    g = torch.Generator().manual_seed(seed)
    w1 = torch.randn(32, 8, generator=g, dtype=torch.float64) * math.sqrt(2.0 / 8)
    w2 = torch.randn(32, 32, generator=g, dtype=torch.float64) * math.sqrt(2.0 / 32)
    parts = [w1.flatten(), torch.zeros(32, dtype=torch.float64), w2.flatten(),
             torch.zeros(32, dtype=torch.float64), torch.zeros(32, dtype=torch.float64),
             torch.zeros(1, dtype=torch.float64)]
    w = torch.cat(parts).to(torch.float32)
    assert w.numel() == PARAMS
    return w.to(device)


def quantize(field: torch.Tensor, bits: int = QUANT_BITS):
    """int32 fixed point q = round(f / scale), scale = max|f| / 2^bits."""
    assert field.dtype == torch.float32 and field.is_contiguous()
    fmax = field.abs().max().item()
    scale = (fmax if fmax > 0 else 1.0) / float(1 << bits)
    q = torch.empty(field.shape, dtype=torch.int32, device=field.device)
    with torch.cuda.device(field.device):
        check(_lib.lib.paliwa_quantize_f32(_ptr(field), _ptr(q), field.numel(),
                                           ctypes.c_float(scale), _stream(field.device)))
    return q, scale


def dequantize(q: torch.Tensor, scale: float) -> torch.Tensor:
    out = torch.empty(q.shape, dtype=torch.float32, device=q.device)
    with torch.cuda.device(q.device):
        check(_lib.lib.paliwa_dequantize_f32(_ptr(q), _ptr(out), q.numel(),
                                             ctypes.c_float(scale), _stream(q.device)))
    return out


def _masks(ndim: int, masks):
    if masks is None:
        masks = [0] * ndim
    return (ctypes.c_uint32 * ndim)(*[int(m) for m in masks])


def forward(q: torch.Tensor, level, lmin, weights, masks=None, update=True,
            stats: torch.Tensor | None = None, overflow: torch.Tensor | None = None,
            axis: int = -1, features: int = FEATURES_DETRENDED):
    """Learned hierarchization of an int32 grid, in place (axes 0..D-1, or
    only `axis`)."""
    D = q.dim()
    with torch.cuda.device(q.device):
        check(_lib.lib.paliwa_learned_hierarchize(
            _ptr(q), D, _lib.i64_array(q.shape), _lib.i32_array(level),
            _lib.i32_array(lmin), int(axis), _ptr(weights), _masks(D, masks), int(update),
            int(features), _ptr(stats), _ptr(overflow), _stream(q.device)))


def inverse(q: torch.Tensor, level, lmin, weights, masks=None, update=True,
            overflow: torch.Tensor | None = None, axis: int = -1,
            features: int = FEATURES_DETRENDED):
    """Inverse of forward, in place (axes D-1..0, or only `axis`)."""
    D = q.dim()
    with torch.cuda.device(q.device):
        check(_lib.lib.paliwa_learned_dehierarchize(
            _ptr(q), D, _lib.i64_array(q.shape), _lib.i32_array(level),
            _lib.i32_array(lmin), int(axis), _ptr(weights), _masks(D, masks), int(update),
            int(features), _ptr(overflow), _stream(q.device)))


def mlp_eval(inputs: torch.Tensor, weights: torch.Tensor) -> torch.Tensor:
    """The network on inputs [n, 8] through the MFMA path of the transform."""
    inputs = inputs.contiguous()
    out = torch.empty(inputs.shape[0], dtype=torch.float32, device=inputs.device)
    with torch.cuda.device(inputs.device):
        check(_lib.lib.paliwa_learned_mlp_eval(_ptr(inputs), _ptr(weights), _ptr(out),
                                               inputs.shape[0], _stream(inputs.device)))
    return out


def _workspace(count: int, device) -> torch.Tensor:
    n = _lib.lib.paliwa_learned_gradient_workspace(int(count))
    return torch.empty(max(1, n), dtype=torch.float32, device=device)


def gradient(pairs: torch.Tensor, weights: torch.Tensor, index: torch.Tensor | None = None):
    """(loss, gradient) of mean |mlp(u) - t| over the rows of pairs [n, >= 9]
    (deterministic: partial sums are added in a fixed order)."""
    pairs = pairs.contiguous()
    grad = torch.zeros(PARAMS, dtype=torch.float32, device=pairs.device)
    loss = torch.zeros(1, dtype=torch.float32, device=pairs.device)
    count = pairs.shape[0] if index is None else index.numel()
    work = _workspace(count, pairs.device)
    with torch.cuda.device(pairs.device):
        check(_lib.lib.paliwa_learned_gradient(
            _ptr(pairs), pairs.shape[1], _ptr(index), count, _ptr(weights),
            _ptr(grad), _ptr(loss), _ptr(work), _stream(pairs.device)))
    return loss, grad


def extract_pairs(grids: Sequence[torch.Tensor], level, lmin, capacity: int,
                  update: bool = True, seed: int = 0,
                  features: int = FEATURES_DETRENDED) -> torch.Tensor:
    """Training pairs {8 network inputs, interpolet residual in correction
    units, clamped to the correction bound} from the interpolet transform of int32
    grids (consumed: they are
    transformed in place), sampled uniformly to about `capacity` pairs (at
    most `capacity`) in a deterministic order."""
    device = grids[0].device
    candidates = 0
    for g in grids:
        for d, l in enumerate(level):
            if l <= lmin[d]:
                continue
            lines = g.numel() // g.shape[d]
            npred = sum((1 << l) >> (level[d] - lev + 1) for lev in range(max(3, lmin[d] + 1), l + 1))
            candidates += lines * npred
    rate = min(1.0, capacity / max(1, candidates))
    threshold = min(0xFFFFFFFF, int(rate * 2.0 ** 32))
    # headroom: the sampled count is binomial around `capacity`
    room = capacity + capacity // 8 + 4096
    pairs = torch.empty(room, 10, dtype=torch.float32, device=device)
    count = torch.zeros(1, dtype=torch.int32, device=device)
    with torch.cuda.device(device):
        for i, g in enumerate(grids):
            check(_lib.lib.paliwa_learned_extract_pairs(
                _ptr(g), g.dim(), _lib.i64_array(g.shape), _lib.i32_array(level),
                _lib.i32_array(lmin), int(update), int(features), _ptr(pairs), _ptr(count),
                room, threshold, seed * 1000 + i, _stream(device)))
    got = int(count.item())
    if got > room:
        raise RuntimeError(f"extract_pairs: {got} sampled pairs overflow the buffer of {room}")
    pairs = pairs[:got]
    # a deterministic order whatever order the workgroups wrote in: the hash
    # key, then the row's contents (equal keys and contents = equal rows)
    words = pairs[:, :10].contiguous().view(torch.int32).to(torch.int64) & 0xFFFFFFFF
    mult = torch.tensor([(2654435761 * (j + 1)) & 0x7FFFFFFF for j in range(9)],
                        dtype=torch.int64, device=device)
    content = (words[:, :9] * mult).sum(1) & 0xFFFFFFFF
    order = torch.argsort((words[:, 9] << 32) | content, stable=True)
    return pairs[order][:capacity].contiguous()


def train(pairs: torch.Tensor, steps: int = 10000, batch: int = 16384, lr: float = 2e-3,
          seed: int = 0, weights: torch.Tensor | None = None, log_every: int = 0):
    """Adam on the mean absolute normalized residual (the detail L1 the
    predict is meant to lower), all in the HIP kernels (gradient, Adam);
    minibatches drawn with a seeded generator."""
    device = pairs.device
    w = init_weights(seed, device) if weights is None else weights.clone()
    m1 = torch.zeros(PARAMS, dtype=torch.float32, device=device)
    m2 = torch.zeros(PARAMS, dtype=torch.float32, device=device)
    grad = torch.zeros(PARAMS, dtype=torch.float32, device=device)
    loss = torch.zeros(1, dtype=torch.float32, device=device)
    gen = torch.Generator(device=device).manual_seed(seed)
    history = []
    n = pairs.shape[0]
    stream = _stream(device)
    work = _workspace(batch, device)
    with torch.cuda.device(device):
        for step in range(1, steps + 1):
            idx = torch.randint(0, n, (batch,), generator=gen, device=device, dtype=torch.int32)
            cur_lr = lr * 0.5 * (1.0 + math.cos(math.pi * (step - 1) / steps))
            check(_lib.lib.paliwa_learned_fill(_ptr(loss), 1, ctypes.c_float(0.0), stream))
            check(_lib.lib.paliwa_learned_gradient(
                _ptr(pairs), pairs.shape[1], _ptr(idx), batch, _ptr(w), _ptr(grad),
                _ptr(loss), _ptr(work), stream))
            check(_lib.lib.paliwa_learned_adam(
                _ptr(w), _ptr(grad), _ptr(m1), _ptr(m2), ctypes.c_float(cur_lr),
                ctypes.c_float(0.9), ctypes.c_float(0.999), ctypes.c_float(1e-8), step,
                stream))
            if log_every and (step % log_every == 0 or step == 1):
                history.append((step, loss.item()))
    return w, history


def choose_masks(q: torch.Tensor, level, lmin, weights, update=True,
                 features: int = FEATURES_DETRENDED, log=None):
    """Per axis, the levels >= 3 where the MLP predict is kept; the interpolet
    is the fallback everywhere else.  The criterion is the detail L1 of the
    whole forward transform, not of the level's own residuals: a nonlinear
    predict along one axis can roughen the data the later axes see, so a
    level that lowers its own residuals may still raise the final L1.  Greedy:
    candidates (axis, level) from the last axis to the first and from the
    finest level to the coarsest, each kept if it lowers the final L1 of the
    masks kept so far.  Returns (masks, [(axis, level, l1, kept), ...])."""
    D = q.dim()
    masks = [0] * D
    work = torch.empty_like(q)

    def final_l1(m):
        work.copy_(q)
        forward(work, level, lmin, weights, m, update, features=features)
        return int(work.abs().sum(dtype=torch.int64).item())

    best = final_l1(masks)
    trail = [(-1, -1, best, True)]
    for d in range(D - 1, -1, -1):
        for lev in range(level[d], max(3, lmin[d] + 1) - 1, -1):
            trial = list(masks)
            trial[d] |= 1 << lev
            l1 = final_l1(trial)
            keep = l1 < best
            trail.append((d, lev, l1, keep))
            if log is not None:
                log(f"  axis {d} level {lev}: L1 {l1} vs {best} -> {'keep' if keep else 'fallback'}")
            if keep:
                masks, best = trial, l1
    del work
    return masks, trail


def detail_stats(coeffs: torch.Tensor, level, lmin, threshold: float):
    """Fraction of detail coefficients with |c| > threshold and their L1 norm
    (float units).  Details: every lattice point except the minimum-level
    lattice; periodic images are excluded (they repeat index 0)."""
    lat = coeffs[tuple(slice(0, 1 << l) for l in level)].double()
    mask = torch.ones(lat.shape, dtype=torch.bool, device=lat.device)
    coarse = mask[tuple(slice(None, None, 1 << (l - m)) for l, m in zip(level, lmin))]
    coarse.fill_(False)
    det = lat[mask].abs()
    return (det > threshold).double().mean().item(), det.sum().item()


def workload(level=(9, 9, 9), train_level=(7, 7, 7), train_seeds=range(100, 132),
             test_seed=4, validation_seed=200, pairs_count=1 << 22, steps=10000,
             batch=16384, lr=2e-3, update=True, features=FEATURES_DETRENDED, device="cuda",
             log=None):
    """The learned-lifting workload: train on turbulent-like fields (64
    Fourier modes with a k^-5/3 spectrum plus 5 tanh fronts of width 0.01),
    choose the per-level masks on a held-out validation field of the test
    size (a seed outside the training and test seeds), and compare, on the
    test field, the learned predict against the interpolet (the same integer
    lifting with the fallback everywhere): fraction of detail coefficients
    above 1e-4 max|f|, detail L1, and the forward+inverse round trip.  Also
    reports masks probed on the test field itself (selection on the test
    data, labelled as such) and the float interpolet of the plain library on
    the same field.  The training loss (the mean absolute residual), length
    and minibatch (10000 steps of 16384 pairs) were chosen on the validation
    field's detail L1 ratio over three training seeds each: 0.602-0.619,
    against 0.861-0.903 with the mean squared residual and 0.605-0.620 /
    0.614-0.630 for 3000 steps of 4096 / 30000 of 16384."""
    import time

    from . import fields
    from . import transform as plain

    def say(msg):
        if log is not None:
            log(msg)

    D = len(level)
    lmin = [0] * D
    # training pairs
    t0 = time.time()
    grids = []
    for seed in train_seeds:
        f = fields.fourier_modes(train_level, modes=64, kmax=32, seed=seed, device=device,
                                 fronts=5, front_width=0.01).to(torch.float32)
        q, _ = quantize(f)
        grids.append(q)
        del f
    pairs = extract_pairs(grids, list(train_level), [0] * D, pairs_count, update,
                          features=features)
    del grids
    torch.cuda.synchronize()
    t_pairs = time.time() - t0
    say(f"pairs: {pairs.shape[0]} in {t_pairs:.1f} s")
    t0 = time.time()
    w0 = init_weights(0, device)
    loss0, _ = gradient(pairs[: 1 << 18], w0)
    weights, hist = train(pairs, steps=steps, batch=batch, lr=lr, log_every=max(1, steps // 10))
    loss1, _ = gradient(pairs[: 1 << 18], weights)
    torch.cuda.synchronize()
    t_train = time.time() - t0
    say(f"train: loss {loss0.item():.4g} -> {loss1.item():.4g} in {t_train:.1f} s")

    # masks on a held-out field of the test size
    fv = fields.fourier_modes(level, modes=64, kmax=32, seed=validation_seed, device=device,
                              fronts=5, front_width=0.01).to(torch.float32)
    qv, _ = quantize(fv)
    del fv
    masks, trail = choose_masks(qv, level, lmin, weights, update, features, log=log)
    del qv
    say(f"masks (validation seed {validation_seed}): {[bin(m) for m in masks]}")

    # test field
    f = fields.fourier_modes(level, modes=64, kmax=32, seed=test_seed, device=device,
                             fronts=5, front_width=0.01).to(torch.float32)
    fmax = f.abs().max().item()
    thr = 1e-4 * fmax
    q, scale = quantize(f)

    def timed(fn):
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1)

    out = {}
    overflow = torch.zeros(1, dtype=torch.int32, device=device)
    for name, m in (("interpolet", [0] * D), ("learned", masks)):
        c = q.clone()
        ms_f = timed(lambda: forward(c, level, lmin, weights, m, update, overflow=overflow,
                                     features=features))
        frac, l1 = detail_stats(c.double() * scale, level, lmin, thr)
        ms_i = timed(lambda: inverse(c, level, lmin, weights, m, update, overflow=overflow,
                                     features=features))
        exact = bool(torch.equal(c, q))
        back = dequantize(c, scale)
        rt = (back - f).abs().max().item() / fmax
        out[name] = {"frac_above_1e-4": frac, "detail_l1": l1, "round_trip_rel": rt,
                     "integer_round_trip_exact": exact, "forward_ms": round(ms_f, 3),
                     "inverse_ms": round(ms_i, 3)}
        del c, back
    # the learned run with the MLP at every level >= 3 (no fallback), for reference
    all_on = [sum(1 << l for l in range(3, level[d] + 1)) for d in range(D)]
    # steady state: the median of 7 runs after 2 (the clock ramps up over the
    # first few back-to-back runs); every run starts from the same grid
    c = torch.empty_like(q)
    runs = []
    for _ in range(9):
        c.copy_(q)
        runs.append(timed(lambda: forward(c, level, lmin, weights, all_on, update,
                                          overflow=overflow, features=features)))
    ms_all = sorted(runs[2:])[3]
    frac, l1 = detail_stats(c.double() * scale, level, lmin, thr)
    # network evaluations of that run: per axis, lines x odd points of levels >= 3
    evals = 0
    for d in range(D):
        lines = q.numel() // q.shape[d]
        evals += lines * sum(1 << (l - 1) for l in range(max(3, lmin[d] + 1), level[d] + 1))
    out["learned_without_fallback"] = {"frac_above_1e-4": frac, "detail_l1": l1,
                                       "forward_ms": round(ms_all, 3),
                                       "forward_ms_runs": [round(t, 3) for t in runs],
                                       "network_evals": evals}
    del c
    # masks probed on the test field itself: selection on the test data, an
    # upper bound on what the masks can give, reported apart
    masks_test, _ = choose_masks(q, level, lmin, weights, update, features)
    c = q.clone()
    forward(c, level, lmin, weights, masks_test, update, features=features)
    frac, l1 = detail_stats(c.double() * scale, level, lmin, thr)
    out["learned_masks_from_test_field"] = {"frac_above_1e-4": frac, "detail_l1": l1,
                                            "masks": masks_test}
    del c
    # the plain float transform of the library (reference arithmetic), hat
    g = f.clone()
    plain.hierarchize(g, level, lmin, level, "hat")
    frac, l1 = detail_stats(g.double(), level, lmin, thr)
    plain.dehierarchize(g, level, lmin, level, "hat")
    out["float_hat"] = {"frac_above_1e-4": frac, "detail_l1": l1,
                        "round_trip_rel": (g - f).abs().max().item() / fmax}
    out["overflow"] = int(overflow.item())
    out["masks"] = masks
    out["mask_field_seed"] = validation_seed
    out["features"] = "detrended" if features == FEATURES_DETRENDED else "raw"
    out["update"] = bool(update)
    out["train"] = {"pairs": int(pairs.shape[0]), "steps": steps, "batch": batch,
                    "loss_before": loss0.item(), "loss_after": loss1.item(),
                    "seconds": round(t_train, 2), "pair_seconds": round(t_pairs, 2),
                    "history": hist}
    out["field"] = {"level": list(level), "points": f.numel(), "max_abs": fmax,
                    "quant_scale": scale}
    li, le = out["interpolet"], out["learned"]
    out["l1_ratio_learned_vs_interpolet"] = le["detail_l1"] / li["detail_l1"]
    out["frac_ratio_learned_vs_interpolet"] = (le["frac_above_1e-4"] / li["frac_above_1e-4"]
                                               if li["frac_above_1e-4"] > 0 else None)
    return out
