#!/usr/bin/env python3
"""Benchmark of the MI355X lifting transform.

Workload: the CDF biorthogonal (predict + update) forward + inverse transform
of one 3-D float64 grid at levels (8, 8, 8) -- 257^3 points including the
periodic image planes (136 MB) -- holding a sum of 16 Fourier modes with
integer wavenumbers 1..32, amplitudes |k|^-5/3 and random phases (seed 1).
One step = paliwa hierarchize + dehierarchize of the whole grid, in place,
already resident in HBM: 6 axis sweeps, each one kernel launch.

value = algorithmic bytes / time: each axis sweep reads and writes every
point once, so a step moves 2 directions * 2 * D * N * 8 bytes.

Also reported: the dominant kernel's achieved bandwidth against the 8 TB/s
HBM3E peak (HIP events around every launch in the timed region), a cold
variant with the 256 MB Infinity Cache flushed before every sweep, and the
CPU baseline: the same transform by the C++/OpenMP restatement of the
reference's schedules on the host cores.  Further sections: float32 storage,
2-D and 6-D workloads per axis, the 1325-grid combination-technique set
(batched vs per-grid, then the combine gather), the learned lifting workload,
a slab's partial-output transform (one shard of a slab decomposition, with its
ghost planes) and raw dump / load throughput.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
USABLE_CPUS = 0  # CPUs this process may use, counted before OpenMP pins threads
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
LEVEL = (8, 8, 8)
WAVELET = "biorthogonal"
METRIC = "FP64 lifting GB/s & %HBM peak (257³); batched CT vs per-grid speedup @1e-12"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--settle-ms", type=float, default=200.0,
                    help="untimed transform work before the warmup steps, so that "
                         "the GPU clocks have ramped up when timing starts")
    ap.add_argument("--no-cpu-baseline", action="store_true",
                    help="skip the host-core baseline (profiling runs)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="wall-time budget of each CPU baseline schedule")
    ap.add_argument("--no-ct", action="store_true",
                    help="skip the combination-technique batch measurement")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-lifting workload")
    ap.add_argument("--ct-only", action="store_true",
                    help="after the timed steps, only the combination-technique section "
                         "(counter runs of the batched plan and the combine)")
    ap.add_argument("--ct-replays", type=int, default=0,
                    help="only build the CT set's batched plan and replay its forward "
                         "direction this many times (kernel traces of the schedule)")
    ap.add_argument("--cold-probe", action="store_true",
                    help="only time the 257^3 sweeps warm, after a 512 MB write flush "
                         "of the Infinity Cache and after a 512 MB read flush (counter runs)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 1 GiB device copy (counter calibration)")
    ap.add_argument("--pitch-probe", action="store_true",
                    help="only time the strided axis-0 sweep of 257 x X x Y float64 grids "
                         "of about the 257^3 grid's size at several row pitches, with the "
                         "fused kernel and with the column tiles")
    return ap.parse_args()


def sweep_kernel(pw, extents, dim, forward):
    """The kernel the automatic schedule runs on axis `dim` of the 257^3
    biorthogonal float64 grid, as the library reports it (paliwa_auto_path):
    the long-line kernel's column tiles on the two strided axes (lines of 256
    points), the fused register kernel on the contiguous one."""
    name = pw.auto_path(dim, extents, LEVEL, [0] * len(LEVEL), WAVELET, forward)
    return "long_ladder column tiles" if name == "long_ladder" and LEVEL[dim] == 8 else name


def copy_probe(torch):
    n = (1 << 30) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8,
                      "launches": 5}))


def cold_probe(torch, reps=5):
    """The six forward/inverse sweeps of the 257^3 biorthogonal grid, each
    timed alone (median of `reps`) in three cache states: warm (the grid was
    just swept), after writing a 512 MB buffer (the Infinity Cache then holds
    dirty lines of another buffer, which the sweep evicts to HBM), and after
    reading a 512 MB buffer (clean lines: eviction costs nothing).  The states
    run in the order warm, write flush, read flush for every sweep, so a
    counter run sees them as consecutive launches of the sweep kernels."""
    import paliwa_amd as pw
    from paliwa_amd import fields

    grid = fields.fourier_modes(LEVEL, modes=16, kmax=32, seed=1, device="cuda")
    level, lmin = list(LEVEL), [0] * len(LEVEL)
    flush = torch.ones(512 * 1024 * 1024 // 4, dtype=torch.float32, device="cuda")
    sink = torch.zeros(1, dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()
    states = {"warm": lambda: None,
              "write_flush": lambda: flush.fill_(1.0),
              "read_flush": lambda: torch.sum(flush, dim=0, out=sink)}
    out = {}
    for fwd in (True, False):
        fn = pw.hierarchize_in if fwd else pw.dehierarchize_in
        for d in range(len(LEVEL)):
            times = {k: [] for k in states}
            fn(d, grid, level, lmin, level, WAVELET)  # warm-up
            for _ in range(reps):
                for name, prep in states.items():
                    if name == "warm":
                        fn(d, grid, level, lmin, level, WAVELET)
                    prep()
                    e0 = torch.cuda.Event(enable_timing=True)
                    e1 = torch.cuda.Event(enable_timing=True)
                    e0.record(stream)
                    fn(d, grid, level, lmin, level, WAVELET)
                    e1.record(stream)
                    torch.cuda.synchronize()
                    times[name].append(e0.elapsed_time(e1))
            key = f"{'fwd' if fwd else 'inv'}_dim{d}"
            out[key] = {k: round(sorted(v)[len(v) // 2], 4) for k, v in times.items()}
    print(json.dumps({"cold_probe_ms": out, "bytes_per_sweep": 2 * grid.numel() * 8}))


def pitch_probe(torch, reps=50):
    """The strided sweep of axis 0 (lines of 257 points, biorthogonal,
    forward and inverse alternating, HIP events over `reps` of each) on
    257 x X x Y float64 grids with X * Y within 6% of 257^2, through two
    explicit schedules: the fused kernel (32-column tiles, 256-byte rows) and
    the long-line kernel's column tiles (64 columns, 512-byte rows).  Each
    keeps its tile shape and bytes per line; only the distance between rows
    (X * Y * 8 bytes) changes.  A tile row starts on a 128-byte line boundary
    in every row only when the pitch is a multiple of 16 points."""
    import ctypes
    import math

    from paliwa_amd import _lib, transform

    stream = torch.cuda.current_stream()

    def level_of(x):
        return int(math.log2(x)) if x & (x - 1) == 0 else int(math.log2(x - 1))

    filters = {True: transform.wavelet_filters("biorthogonal"),
               False: transform.wavelet_filters("biorthogonal", reconstruct=True)}
    out = []
    shapes = ((256, 256), (2049, 32), (1025, 64), (513, 128), (512, 129), (257, 256),
              (257, 257), (129, 513), (65, 1025), (33, 2049))
    # the list twice, the first pass untimed in the report (clocks settle)
    for k, (X, Y) in enumerate(shapes + shapes):
        shape = (257, X, Y)
        level = [8, level_of(X), level_of(Y)]
        g = torch.randn(shape, dtype=torch.float64, device="cuda")
        row = {"shape": list(shape), "pitch_bytes": X * Y * 8, "pitch_mod_128": X * Y * 8 % 128}
        for path in ("fused", "long_ladder"):
            calls = []
            for fwd in (True, False):
                levels = list(range(8, 0, -1)) if fwd else list(range(1, 9))
                f = filters[fwd]
                calls.append((ctypes.c_void_p(g.data_ptr()), _lib.PALIWA_FLOAT64, 3,
                              _lib.i64_array(shape), _lib.i32_array(level), _lib.i32_array(level),
                              0, _lib.i32_array(levels), len(levels), _lib.filter_array(f), len(f),
                              _lib.PATHS[path], ctypes.c_void_p(stream.cuda_stream)))
            fn = _lib.lib.paliwa_transform_in_path
            for _ in range(3):
                for a in calls:
                    assert fn(*a) == 0, _lib.last_error()
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(reps):
                for a in calls:
                    fn(*a)
            e1.record(stream)
            torch.cuda.synchronize()
            us = e0.elapsed_time(e1) / (2 * reps) * 1e3
            name = "fused" if path == "fused" else "columns"
            row[f"{name}_us"] = round(us, 2)
            row[f"{name}_GBps"] = round(2 * g.numel() * 8 / (us * 1e-6) / 1e9, 1)
        if k >= len(shapes):
            out.append(row)
        del g
    print(json.dumps({"pitch_probe": out}))


def partial_benchmark(torch, ranks=8, reps=20):
    """One rank's share of a 256^3 float64 grid cut into `ranks` slabs along
    axis 0 (the shard of a slab decomposition): the partial-output forward
    transform of axis 0 on the slab and its ghost planes (the step after a
    ghost exchange), for three dependency depths -- hat (one ghost plane),
    biorthogonal down to level 4, biorthogonal down to level 0 (whose
    dependency cone is the whole axis) -- beside the whole-grid axis-0 sweep."""
    import paliwa_amd as pw
    from paliwa_amd.partial import PartialTransform

    L = 8
    P = 1 << L
    n = P // ranks
    plane = P * P * 8
    block = torch.randn(n, P, P, dtype=torch.float64, device="cuda")
    full = torch.randn(P, P, P, dtype=torch.float64, device="cuda")

    def timed(fn):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    cases = {}
    for name, lmin in (("hat", 0), ("biorthogonal", 4), ("biorthogonal", 0)):
        plan = PartialTransform([n, P, P], 0, L, list(range(L, lmin, -1)),
                                pw.wavelet_filters(name), local_start=0,
                                output_start=0, output_count=n)
        g = len(plan.ghost_indices)
        ghosts = torch.randn(max(g, 1), P, P, dtype=torch.float64, device="cuda")[:g]
        ms = timed(lambda: plan(block, ghosts if g else None))
        full_ms = timed(lambda: pw.hierarchize_in(0, full, [L] * 3, [lmin] * 3, [L] * 3, name))
        moved = (n + g) * plane + n * plane  # inputs read, outputs written
        cases[f"{name}_lmin{lmin}"] = {
            "ghost_planes": g,
            "ms": round(ms, 4),
            "GBps": round(moved / (ms * 1e-3) / 1e9, 1),
            "full_axis_sweep_ms": round(full_ms, 4),
            "full_axis_sweep_GBps": round(2 * P ** 3 * 8 / (full_ms * 1e-3) / 1e9, 1),
        }
        plan.close()
    return {"workload": f"axis 0 of a 256^3 float64 grid in {ranks} slabs: slab 0 "
                        f"({n} planes) and its ghost planes, forward transform",
            "cases": cases}


def rawio_benchmark(torch):
    """Raw dump and load (the reference's dump_chunk_span_to_binary_file
    format) of a 257^3 float64 grid through a temporary file: device-host copy
    plus file I/O, bounded by PCIe and the host's file system, not by HBM."""
    import tempfile
    import time

    from paliwa_amd import rawio

    grid = torch.randn(257, 257, 257, dtype=torch.float64, device="cuda")
    nbytes = grid.numel() * 8
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "full_grid_8_8_8.raw")
        rawio.dump(grid, path)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rawio.dump(grid, path)
        torch.cuda.synchronize()
        dump_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        back = rawio.load(path, grid.shape, torch.float64)
        torch.cuda.synchronize()
        load_s = time.perf_counter() - t0
        same = bool(torch.equal(back, grid))
    return {"workload": "257^3 float64 raw dump and load (temporary file)",
            "bytes": nbytes, "dump_GBps": round(nbytes / dump_s / 1e9, 2),
            "load_GBps": round(nbytes / load_s / 1e9, 2), "round_trip_equal": same}


def cpu_baseline(grid_cpu, budget_s):
    """Time the oracle's OpenMP schedules on the host cores (bounded sample)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import oracle_bridge as ob

    level = list(LEVEL)
    lmin = [0, 0, 0]
    n_points = grid_cpu.size
    bytes_per_rep = 2 * 2 * len(level) * n_points * 8
    best = None
    detail = []
    for label, sched in (("packed-poles", ob.OMP_PACKED_POLES), ("direct", ob.OMP_DIRECT)):
        a = np.ascontiguousarray(grid_cpu.copy())
        ob.hierarchize(a, level, lmin, level, WAVELET, sched)  # warm-up
        ob.dehierarchize(a, level, lmin, level, WAVELET, sched)
        times, t0 = [], time.perf_counter()
        while len(times) < 5 or time.perf_counter() - t0 < budget_s:
            t1 = time.perf_counter()
            ob.hierarchize(a, level, lmin, level, WAVELET, sched)
            ob.dehierarchize(a, level, lmin, level, WAVELET, sched)
            times.append(time.perf_counter() - t1)
        dt = float(np.median(times))
        gbs = bytes_per_rep / dt / 1e9
        detail.append(f"{label} median of {len(times)} reps {dt:.3f} s/rep {gbs:.2f} GB/s")
        if best is None or gbs > best:
            best = gbs
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = USABLE_CPUS or len(os.sched_getaffinity(0))
    return {
        "value": round(best, 3),
        "unit": "GB/s",
        "cores": ob.num_threads(),
        "kind": "port",
        "sample": "257^3 biorthogonal float64 forward+inverse, C++/OpenMP "
                  "restatement of the reference's CPU schedules (oracle/), "
                  f"{ob.num_threads()} threads (OMP_PROC_BIND="
                  f"{os.environ.get('OMP_PROC_BIND')}, OMP_PLACES="
                  f"{os.environ.get('OMP_PLACES')}) of the {usable} CPUs this process may "
                  f"use ({os.cpu_count()} in the host) on {model}; median per step after a "
                  "warm-up: " + "; ".join(detail),
    }


def fp32_variant(torch, pw, original64, steps, warmup):
    """The same workload with float32 storage and arithmetic: throughput and
    error against float64 (max-rel and L2 of the coefficients, mass = the
    coarse coefficient times the lattice size against the float64 lattice
    sum, round trip)."""
    level, lmin = list(LEVEL), [0] * len(LEVEL)
    g32 = original64.to(torch.float32)
    stream = torch.cuda.current_stream()

    def step():
        pw.hierarchize(g32, level, lmin, level, WAVELET)
        pw.dehierarchize(g32, level, lmin, level, WAVELET)

    for _ in range(warmup):
        step()
    g32.copy_(original64)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    rt = (g32.double() - original64).abs().max().item() / original64.abs().max().item()
    # coefficient errors against float64
    c64 = original64.clone()
    pw.hierarchize(c64, level, lmin, level, WAVELET)
    c32 = original64.to(torch.float32)
    pw.hierarchize(c32, level, lmin, level, WAVELET)
    torch.cuda.synchronize()
    d = c32.double() - c64
    lat = original64[tuple(slice(0, 1 << l) for l in level)]
    n_lat = lat.numel()
    mass64 = lat.sum().item()
    mass_rel = abs(c32[0, 0, 0].item() * n_lat - mass64) / lat.abs().sum().item()
    n = original64.numel()
    return {
        "GBps": round(2 * 2 * len(level) * n * 4 / (ms * 1e-3) / 1e9, 1),
        "ms_per_step": round(ms, 4),
        "max_rel_err_vs_f64": d.abs().max().item() / c64.abs().max().item(),
        "l2_rel_err_vs_f64": (d.norm() / c64.norm()).item(),
        "mass_rel_err": mass_rel,
        "round_trip_rel_err": rt,
    }


def ct_benchmark(torch, pw, fields, reps=5):
    """Combination-technique set: 1325 4-D float64 grids, |l|_1 in 13..16,
    l_i >= 1, hat basis, periodic images; cosine-product fields (seed 2).
    Batched plan (one graph) vs one hierarchize call per grid (5300 kernel
    launches), both timed with events on the same stream."""
    import ctypes
    import math

    import numpy as np

    from paliwa_amd import _lib

    levels = pw.combination_levels(4, (13, 14, 15, 16), 1)
    rng = np.random.default_rng(2)
    lmin, lmax = [0] * 4, [13] * 4
    grids = []
    for l in levels:
        a = rng.uniform(1.0, 4.0, size=4)
        phi = rng.uniform(0.0, 2 * math.pi, size=4)
        grids.append(fields.cosine_product(l, a, phi, device="cuda"))
    initial = [g.clone() for g in grids]
    points = sum(g.numel() for g in grids)
    stream = torch.cuda.current_stream()
    batch = pw.CombinationBatch(grids, levels, lmin, lmax)

    def timed(fn, n):
        fn()  # warm-up (graph capture, kernel attributes)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(n):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / n

    rule_ms = timed(lambda: batch.hierarchize("hat"), reps)
    # the plan's autotuner: box schedule, tile widths and stream schedule
    # timed on these grids (restored afterwards), no cache so every run tunes
    # afresh
    tune = batch.tune("hat", forward=True, force=True, measure=True)
    batched_ms = timed(lambda: batch.hierarchize("hat"), reps)

    # per grid: prebuilt C-ABI arguments, a tight loop of hierarchize calls
    fn = _lib.lib.paliwa_hierarchize
    h = ctypes.c_void_p(stream.cuda_stream)
    mn, mx = _lib.i32_array(lmin), _lib.i32_array(lmax)
    args = [(ctypes.c_void_p(g.data_ptr()), _lib.PALIWA_FLOAT64, 4,
             _lib.i64_array(g.shape), _lib.i32_array(l), mn, mx, b"hat", h)
            for g, l in zip(grids, levels)]

    def per_grid():
        for a in args:
            fn(*a)

    per_grid_ms = timed(per_grid, max(1, reps // 2))

    # equality of the two forms on the same input
    for g, i in zip(grids, initial):
        g.copy_(i)
    batch.hierarchize("hat")
    out_batched = [g.clone() for g in grids]
    for g, i in zip(grids, initial):
        g.copy_(i)
    per_grid()
    torch.cuda.synchronize()
    max_rel = 0.0
    for a, b in zip(out_batched, grids):
        d = (a - b).abs().max().item()
        s = b.abs().max().item()
        max_rel = max(max_rel, d / s if s > 0 else d)
    info = batch.info()
    boxes = batch.boxes()
    batch.close()
    bytes_fwd = 2 * 4 * points * 8
    combine = combine_benchmark(torch, levels, out_batched, timed, reps)
    cpp = cpp_per_grid_loop()
    return {
        "workload": "1325 4-D float64 grids, |l|_1 in {13..16}, l_i >= 1, hat, "
                    "forward, periodic images",
        "grids": len(levels),
        "points": points,
        "batched_ms": round(batched_ms, 4),
        "rule_ms": round(rule_ms, 4),
        "tuned_ms": round(batched_ms, 4),
        "tuner": {"rule_ms_median": round(tune["rule_ms"], 4),
                  "tuned_ms_median": round(tune["tuned_ms"], 4),
                  "evaluations": tune["evaluations"], "streams": tune["streams"]},
        "per_grid_ms": round(per_grid_ms, 4),
        "speedup": round(per_grid_ms / batched_ms, 2),
        # the metric's bytes: one read and one write of every point per axis
        "algorithmic_GBps": round(bytes_fwd / (batched_ms * 1e-3) / 1e9, 1),
        # what the plan moves: one read and write per segment (box or axis)
        "plan_bytes": 2 * info["points_read_per_direction"] * 8,
        "plan_GBps": round(2 * info["points_read_per_direction"] * 8
                           / (batched_ms * 1e-3) / 1e9, 1),
        "boxes": list(boxes),
        "batched_launches": info["launches_per_direction"],
        "per_grid_launches": 4 * len(levels),
        "max_rel_diff_batched_vs_per_grid": max_rel,
        "cpp_per_grid": cpp,
        "combine": combine,
    }


def ct_set_grids(torch, fields):
    """The CT set of the bench: 1325 4-D float64 grids with cosine-product
    fields (seed 2)."""
    import math

    import numpy as np
    import paliwa_amd as pw

    levels = pw.combination_levels(4, (13, 14, 15, 16), 1)
    rng = np.random.default_rng(2)
    grids = []
    for l in levels:
        a = rng.uniform(1.0, 4.0, size=4)
        phi = rng.uniform(0.0, 2 * math.pi, size=4)
        grids.append(fields.cosine_product(l, a, phi, device="cuda"))
    return levels, grids


def ct_replays(torch, count):
    """The batched plan of the CT set (rule-based settings), forward
    direction replayed `count` times after two warm-up replays."""
    import paliwa_amd as pw
    from paliwa_amd import fields

    levels, grids = ct_set_grids(torch, fields)
    batch = pw.CombinationBatch(grids, levels, [0] * 4, [13] * 4)
    for _ in range(2):
        batch.hierarchize("hat")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(count):
        batch.hierarchize("hat")
    torch.cuda.synchronize()
    print(json.dumps({"ct_replays": count,
                      "ms_per_replay": round((time.perf_counter() - t0) * 1e3 / count, 4),
                      "launches_per_replay": batch.info()["launches_per_direction"]}))
    batch.close()


def cpp_per_grid_loop():
    """The reference's per-grid loop in C++ through the drop-in header
    (tests/cpp/ct_per_grid.cpp): one stream, and 32 streams round-robin as the
    reference's partitioned instances, beside the batched plan on the same
    grids (no autotuning there)."""
    import json
    import subprocess

    exe = os.path.join(ROOT, "tests", "cpp", "build", "ct_per_grid")
    if not os.path.exists(exe):
        return {"skipped": "tests/cpp/build/ct_per_grid is not built"}
    out = subprocess.run([exe, "3"], capture_output=True, text=True, timeout=240)
    if out.returncode != 0:
        return {"error": (out.stdout + out.stderr)[-400:]}
    return json.loads(out.stdout.strip().splitlines()[-1])


def per_axis_benchmark(torch, grid, level, wavelet, reps=20):
    """Per-axis sweep times of one grid (forward and inverse), each sweep
    timed over `reps` back-to-back launches with prebuilt C-ABI arguments
    (HIP events on the launch stream); algorithmic GB/s = 2 * N * s / t."""
    import ctypes

    from paliwa_amd import _lib

    D = len(level)
    stream = torch.cuda.current_stream()
    dtype = _lib.PALIWA_FLOAT64 if grid.dtype == torch.float64 else _lib.PALIWA_FLOAT32
    nbytes = 2 * grid.numel() * grid.element_size()
    out = {}
    for fwd in (True, False):
        fn = _lib.lib.paliwa_hierarchize_in if fwd else _lib.lib.paliwa_dehierarchize_in
        for d in range(D):
            args = (ctypes.c_void_p(grid.data_ptr()), dtype, D, _lib.i64_array(grid.shape),
                    _lib.i32_array(level), _lib.i32_array([0] * D), _lib.i32_array(level), d,
                    wavelet.encode(), ctypes.c_void_p(stream.cuda_stream))
            for _ in range(3):
                assert fn(*args) == 0, _lib.last_error()
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(reps):
                fn(*args)
            e1.record(stream)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            out[f"{'fwd' if fwd else 'inv'}_dim{d}"] = {
                "us": round(ms * 1e3, 2), "GBps": round(nbytes / (ms * 1e-3) / 1e9, 1)}
    total_us = sum(v["us"] for v in out.values())
    return {"per_axis": out, "fwd_inv_us": round(total_us, 2),
            "GBps": round(2 * D * nbytes / (total_us * 1e-6) / 1e9, 1)}


def other_workloads(torch, fields):
    """The 2-D hat grid (1025^2), the anisotropic 6-D grid (5,5,4,4,3,3)
    with both wavelets and the 513^3 float64 biorthogonal grid (larger than
    the Infinity Cache): per-axis GB/s.  The first two are bit-identical to
    the CPU oracle (tests/test_workloads_gpu.py); the 513^3 grid runs the same
    kernels at period 512 (tests/test_parity_gpu.py covers every period)."""
    out = {}
    g = fields.gaussian_bump_2d((10, 10), seed=0, device="cuda")
    out["hat_2d_1025"] = per_axis_benchmark(torch, g, [10, 10], "hat")
    g = fields.cosine_modes_noise((5, 5, 4, 4, 3, 3), seed=3, device="cuda")
    for wavelet in ("hat", "biorthogonal"):
        out[f"{wavelet}_6d_33x33x17x17x9x9"] = per_axis_benchmark(
            torch, g, [5, 5, 4, 4, 3, 3], wavelet)
    del g
    # long lines (4097 points per axis): the long-line kernel on both axes
    g = fields.gaussian_bump_2d((12, 12), seed=0, device="cuda")
    out["biorthogonal_2d_4097"] = per_axis_benchmark(torch, g, [12, 12], "biorthogonal")
    del g
    # a grid four times the Infinity Cache (1.08 GB): every sweep streams HBM
    g = fields.fourier_modes((9, 9, 9), modes=16, kmax=32, seed=1, device="cuda")
    out["biorthogonal_3d_513"] = per_axis_benchmark(torch, g, [9, 9, 9], "biorthogonal")
    del g
    return out


def combine_benchmark(torch, levels, hierarchized, timed, reps):
    """The step after the batched hierarchize: collect the 1325 hierarchized
    grids onto sparse-grid storage (coefficients +1, -3, +3, -1 for
    |l|_1 = 16, 15, 14, 13; subspaces in more than one grid)."""
    from paliwa_amd.combine import SparseGridCombination

    coef_of = {16: 1.0, 15: -3.0, 14: 3.0, 13: -1.0}
    comb = SparseGridCombination(hierarchized, levels, [coef_of[sum(l)] for l in levels],
                                 [13] * 4)
    info = comb.info()
    sparse = torch.empty(info["sparse_points"], dtype=torch.float64, device="cuda")
    ms = timed(lambda: comb.gather(sparse), reps)
    comb.close()
    nbytes = (info["terms"] + info["sparse_points"]) * 8
    return {
        "subspaces": info["subspaces"],
        "sparse_points": info["sparse_points"],
        "terms": info["terms"],
        "gather_ms": round(ms, 4),
        "gather_GBps": round(nbytes / (ms * 1e-3) / 1e9, 1),
        "algorithmic_bytes": nbytes,
    }


def network_rate(run):
    flop = 2624 * run["network_evals"]
    tflops = flop / (run["forward_ms"] * 1e-3) / 1e12
    return {"evaluations": run["network_evals"], "flop_per_evaluation": 2624,
            "all_levels_forward_ms": run["forward_ms"], "runs_ms": run["forward_ms_runs"],
            "timing": "median of 7 back-to-back runs after 2", "TFLOPs": round(tflops, 1),
            "bf16_dense_peak_TFLOPs": 2500.0, "frac": round(tflops / 2500.0, 4),
            "bound": "valu"}


def learned_benchmark():
    """The learned-lifting workload: 513^3 float32 test field (seed 4),
    training on 32 fields of 129^3 (seeds 100-131), 2^22 pairs; the learned
    predictor with the interpolet fallback against the interpolet alone."""
    from paliwa_amd import learned

    r = learned.workload()
    li, le = r["interpolet"], r["learned"]
    pts = r["field"]["points"]
    tt = r["learned_masks_from_test_field"]
    # the input spec taken literally: the 8 neighbours over their max-abs
    raw = learned.workload(features=learned.FEATURES_RAW)
    return {
        "workload": "learned lifting, 3-D float32 (9,9,9) = 513^3, detrended inputs, "
                    "update kept, interpolet fallback; masks chosen on a held-out field "
                    f"(seed {r['mask_field_seed']}), results on the test field (seed 4)",
        "detail_l1": {"learned": le["detail_l1"], "interpolet": li["detail_l1"],
                      "ratio": round(r["l1_ratio_learned_vs_interpolet"], 4),
                      "float_hat": r["float_hat"]["detail_l1"]},
        "frac_above_1e-4": {"learned": le["frac_above_1e-4"], "interpolet": li["frac_above_1e-4"],
                            "ratio": round(r["frac_ratio_learned_vs_interpolet"], 4),
                            "float_hat": r["float_hat"]["frac_above_1e-4"]},
        "masks_from_test_field": {
            "detail_l1_ratio": round(tt["detail_l1"] / li["detail_l1"], 4),
            "frac_ratio": round(tt["frac_above_1e-4"] / li["frac_above_1e-4"], 4),
            "masks": tt["masks"]},
        "literal_input_spec": {
            "features": "raw neighbours over their max-abs",
            "detail_l1_ratio": round(raw["l1_ratio_learned_vs_interpolet"], 4),
            "frac_ratio": round(raw["frac_ratio_learned_vs_interpolet"], 4),
            "masks": raw["masks"],
            "loss_before": raw["train"]["loss_before"],
            "loss_after": raw["train"]["loss_after"]},
        "round_trip_rel": le["round_trip_rel"],
        "integer_round_trip_exact": le["integer_round_trip_exact"],
        "masks": r["masks"],
        "forward_ms": le["forward_ms"],
        "inverse_ms": le["inverse_ms"],
        "interpolet_forward_ms": li["forward_ms"],
        "forward_Gpoints_per_s": round(pts / (le["forward_ms"] * 1e-3) / 1e9, 2),
        # the network at every level >= 3 of every axis: its arithmetic against
        # the bf16 dense MFMA peak (2 * (8*32 + 32*32 + 32) = 2624 flop per
        # evaluation); the sweep is bound by the VALU (the packed-fp16 GELU
        # gate and the bf16 conversions of the activations), not by the matrix
        # cores
        "network": network_rate(r["learned_without_fallback"]),
        "train": {k: r["train"][k] for k in ("pairs", "steps", "batch", "loss_before",
                                             "loss_after", "seconds")},
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: only one GPU is supported here (asked for {args.gpus}); "
              "the transform of one grid runs on one GPU and nothing is exchanged "
              "between GPUs.", file=sys.stderr)
        sys.exit(2)
    # the CPU baseline's OpenMP threads: one per core, packed (set before the
    # oracle's OpenMP runtime starts, which then pins the main thread; the
    # CPUs this process may use are counted first)
    global USABLE_CPUS
    USABLE_CPUS = len(os.sched_getaffinity(0))
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    import torch

    if not torch.cuda.is_available():
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)
    if args.copy_probe:
        copy_probe(torch)
        return
    if args.ct_replays:
        ct_replays(torch, args.ct_replays)
        return
    if args.cold_probe:
        cold_probe(torch)
        return
    if args.pitch_probe:
        pitch_probe(torch)
        return

    import paliwa_amd as pw
    from paliwa_amd import fields

    grid = fields.fourier_modes(LEVEL, modes=16, kmax=32, seed=1, device="cuda")
    original = grid.clone()
    n_points = grid.numel()
    D = len(LEVEL)
    level, lmin = list(LEVEL), [0] * D
    bytes_per_sweep = 2 * n_points * 8
    bytes_per_step = 2 * D * bytes_per_sweep
    stream = torch.cuda.current_stream()

    # per-axis C-ABI calls with prebuilt arguments, for the per-launch event
    # pass: cheap enough on the host that the GPU queue never runs dry
    import ctypes

    from paliwa_amd import _lib

    axis_calls = []
    for fwd in (True, False):
        fn = _lib.lib.paliwa_hierarchize_in if fwd else _lib.lib.paliwa_dehierarchize_in
        for d in range(D):
            axis_calls.append((("fwd" if fwd else "inv", d), fn, (
                ctypes.c_void_p(grid.data_ptr()), _lib.PALIWA_FLOAT64, D,
                _lib.i64_array(grid.shape), _lib.i32_array(level), _lib.i32_array(lmin),
                _lib.i32_array(level), d, WAVELET.encode(), ctypes.c_void_p(stream.cuda_stream))))

    def step(events=None):
        if events is None:
            # the reference's entry points: all axes of one direction per call
            pw.hierarchize(grid, level, lmin, level, WAVELET)
            pw.dehierarchize(grid, level, lmin, level, WAVELET)
            return
        for key, fn, call_args in axis_calls:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            rc = fn(*call_args)
            e1.record(stream)
            if rc != 0:
                raise RuntimeError(_lib.last_error())
            events.append((key, e0, e1))

    # clock settle (untimed), then the warmup steps
    settle_t0 = time.perf_counter()
    settle_rounds = 0
    while (time.perf_counter() - settle_t0) * 1e3 < args.settle_ms:
        for _ in range(10):
            step()
        torch.cuda.synchronize()
        settle_rounds += 1
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    # timed region: K steps, nothing but the transform on the stream
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_start.record(stream)
    for _ in range(args.steps):
        step()
    t_end.record(stream)
    torch.cuda.synchronize()
    total_ms = t_start.elapsed_time(t_end)
    ms_per_step = total_ms / args.steps

    # the same K steps again with a HIP event pair around every launch:
    # per-kernel durations for the roofline
    events = []
    for _ in range(3):  # slack in the queue before the first recorded launch
        step()
    for _ in range(args.steps):
        step(events)
    torch.cuda.synchronize()

    per_kernel = {}
    for key, e0, e1 in events:
        per_kernel.setdefault(key, []).append(e0.elapsed_time(e1))
    avg = {k: sum(v) / len(v) for k, v in per_kernel.items()}
    dominant = max(avg, key=avg.get)
    dom_ms = avg[dominant]
    achieved = bytes_per_sweep / (dom_ms * 1e-3) / 1e9

    # cold: the Infinity Cache holds another 512 MB buffer when the sweep
    # starts -- read before it (clean lines: the grid comes from HBM) or
    # written before it (dirty lines, which the sweep's own traffic evicts to
    # HBM on top of its reads and writes)
    flush = torch.ones(512 * 1024 * 1024 // 4, dtype=torch.float32, device="cuda")
    sink = torch.zeros(1, dtype=torch.float32, device="cuda")
    cold, cold_dirty = {}, {}
    for fwd in (True, False):
        fn = pw.hierarchize_in if fwd else pw.dehierarchize_in
        for d in range(D):
            times = {"clean": [], "dirty": []}
            for _ in range(3):
                for kind in ("clean", "dirty"):
                    if kind == "clean":
                        torch.sum(flush, dim=0, out=sink)
                    else:
                        flush.fill_(1.0)
                    e0 = torch.cuda.Event(enable_timing=True)
                    e1 = torch.cuda.Event(enable_timing=True)
                    e0.record(stream)
                    fn(d, grid, level, lmin, level, WAVELET)
                    e1.record(stream)
                    torch.cuda.synchronize()
                    times[kind].append(e0.elapsed_time(e1))
            cold[("fwd" if fwd else "inv", d)] = sorted(times["clean"])[1]
            cold_dirty[("fwd" if fwd else "inv", d)] = sorted(times["dirty"])[1]
    cold_step_ms = sum(cold.values())
    cold_dirty_step_ms = sum(cold_dirty.values())
    del flush

    # practical streaming ceiling for this footprint: an in-place elementwise
    # pass (read + write every element once) over a buffer of the grid's size
    scratch = torch.rand(n_points, dtype=torch.float64, device="cuda")
    for _ in range(5):
        scratch.mul_(1.0)
    torch.cuda.synchronize()
    c0 = torch.cuda.Event(enable_timing=True)
    c1 = torch.cuda.Event(enable_timing=True)
    c0.record(stream)
    for _ in range(50):
        scratch.mul_(1.0)
    c1.record(stream)
    torch.cuda.synchronize()
    stream_ms = c0.elapsed_time(c1) / 50
    del scratch

    # round-trip error after every transform above (warmup, timed and
    # per-launch steps, the cold sweeps), and after one forward + inverse of
    # the original data
    torch.cuda.synchronize()
    scale = original.abs().max().item()
    rt_err = (grid - original).abs().max().item() / scale
    once = original.clone()
    pw.hierarchize(once, level, lmin, level, WAVELET)
    pw.dehierarchize(once, level, lmin, level, WAVELET)
    torch.cuda.synchronize()
    rt_once = (once - original).abs().max().item() / scale
    del once
    n_round_trips = 10 * settle_rounds + args.warmup + 2 * args.steps + 6

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get("bytes_per_launch")
        except Exception:
            traffic = None

    record = {
        "metric": METRIC,
        "value": round(bytes_per_step / (ms_per_step * 1e-3) / 1e9, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (16 Fourier modes, |k|^-5/3, seed 1)",
        "config": {
            "workload": "biorthogonal forward+inverse, one 3-D grid, levels (8,8,8), "
                        "257^3 points with periodic images",
            "points": n_points,
            "bytes_per_step": bytes_per_step,
            "hbm_fraction_of_step": round(bytes_per_step / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "settle_ms": args.settle_ms,
            "inplace_stream_GBps_same_footprint": round(bytes_per_sweep / (stream_ms * 1e-3) / 1e9, 1),
            "cold_ms_per_step": round(cold_step_ms, 4),
            "cold_GBps": round(bytes_per_step / (cold_step_ms * 1e-3) / 1e9, 2),
            "per_kernel_ms": {f"{k[0]}_dim{k[1]}": round(v, 4) for k, v in sorted(avg.items())},
            "cold_per_kernel_ms": {f"{k[0]}_dim{k[1]}": round(v, 4) for k, v in sorted(cold.items())},
            "cold_dirty_ms_per_step": round(cold_dirty_step_ms, 4),
            "cold_dirty_per_kernel_ms": {f"{k[0]}_dim{k[1]}": round(v, 4)
                                         for k, v in sorted(cold_dirty.items())},
            "round_trip_rel_err": rt_once,
            "accumulated_rel_err": {"round_trips": n_round_trips, "max_rel": rt_err},
        },
        "roofline": {
            "bound": "hbm",
            "kernel": f"{sweep_kernel(pw, list(grid.shape), dominant[1], dominant[0] == 'fwd')} "
                      f"{dominant[0]} dim{dominant[1]}",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "residency": "warm: the 136 MB grid stays in the 256 MB Infinity Cache "
                         "between sweeps; cold: the cache holds another 512 MB buffer, "
                         "read before the sweep (clean); cold_dirty: written before it, "
                         "so the sweep also evicts dirty lines to HBM",
            "cold": {
                "achieved": round(bytes_per_sweep / (cold[dominant] * 1e-3) / 1e9, 2),
                "frac": round(bytes_per_sweep / (cold[dominant] * 1e-3) / 1e9
                              / HBM_PEAK_GBS, 4),
            },
            "cold_dirty": {
                "achieved": round(bytes_per_sweep / (cold_dirty[dominant] * 1e-3) / 1e9, 2),
                "frac": round(bytes_per_sweep / (cold_dirty[dominant] * 1e-3) / 1e9
                              / HBM_PEAK_GBS, 4),
            },
        },
        "cpu_baseline": None,
    }
    if args.ct_only:
        record["ct"] = ct_benchmark(torch, pw, fields)
        print(json.dumps(record))
        return
    record["fp32"] = fp32_variant(torch, pw, original, args.steps, args.warmup)
    record["other_workloads"] = other_workloads(torch, fields)
    # the same kernel on a grid four times the Infinity Cache (513^3 float64,
    # 1.08 GB): its slowest sweep, every byte from HBM
    big = record["other_workloads"]["biorthogonal_3d_513"]["per_axis"]
    slow_key = max(big, key=lambda k: big[k]["us"])
    record["roofline"]["beyond_cache_513"] = {
        "kernel": f"fused_ladder {slow_key.replace('_', ' ')}",
        "achieved": big[slow_key]["GBps"],
        "frac": round(big[slow_key]["GBps"] / HBM_PEAK_GBS, 4),
    }
    if not args.no_ct:
        record["ct"] = ct_benchmark(torch, pw, fields)
    if not args.no_learned:
        record["learned"] = learned_benchmark()
    record["partial"] = partial_benchmark(torch)
    record["rawio"] = rawio_benchmark(torch)
    if not args.no_cpu_baseline:
        record["cpu_baseline"] = cpu_baseline(original.cpu().numpy(), args.cpu_seconds)
    print(json.dumps(record), flush=True)


if __name__ == "__main__":
    main()
