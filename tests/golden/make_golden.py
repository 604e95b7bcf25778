#!/usr/bin/env python3
"""Writes the golden vectors under tests/golden/.

1. plan_multipass.json -- the value known-answer test of the reference,
   test/test_paliwa_plan.cpp:114-197 (TEST transform_plan.multipass_sparse_values):
   a periodic 64-point line (maximum level 6) holding 1 + sin(0.7 i), a grid
   stored at level 4 (every 4th point), filter lists of 1..5 passes with every
   parity mask and the coefficients {-.5, 1, .125} / {0, 1, .25}, and the level
   orders {4,3,2,1} and {1,2,3,4}.  The expected values come from the test's
   own brute-force loop (lines 128-139), restated below in finest coordinates
   exactly as the test writes it; Python evaluates f0*a + f1*b + f2*c left to
   right in IEEE double without contraction, as the C++ does.  Floats are
   stored as hex strings so the comparison can be bit-exact.

2. regression_*.raw + regression.json -- small grids (inputs and outputs) of
   the CPU oracle, in the reference's raw dump format (headerless, row-major,
   native-endian; paliwa/paliwa_io.hpp:13-23).  They pin the oracle against
   accidental change and give the GPU tests fixed vectors.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def plan_multipass_cases():
    maximum, level, spacing, period = 6, 4, 4, 64
    cases = []
    for passes in range(1, 6):
        for mask in range(1 << passes):
            filters = []
            for p in range(passes):
                coef = (0.0, 1.0, 0.25) if p % 2 else (-0.5, 1.0, 0.125)
                filters.append((1 if mask & (1 << p) else 0, coef))
            for levels in ([4, 3, 2, 1], [1, 2, 3, 4]):
                inp = [1.0 + math.sin(i * 0.7) for i in range(period)]
                ref = list(inp)
                for l in levels:
                    stride = 1 << (maximum - l)
                    for offset, f in filters:
                        for i in range(offset * stride, period, 2 * stride):
                            ref[i] = (f[0] * ref[(i + period - stride) % period]
                                      + f[1] * ref[i]
                                      + f[2] * ref[(i + stride) % period])
                cases.append({
                    "passes": passes,
                    "mask": mask,
                    "filters": [[o, list(c)] for o, c in filters],
                    "levels": levels,
                    "level": level,
                    "maximum_level": maximum,
                    "input": [inp[k * spacing].hex() for k in range(period // spacing)],
                    "expected": [ref[k * spacing].hex() for k in range(period // spacing)],
                })
    return cases


REGRESSION = [
    # name, dtype, level, extents, minimum level, wavelet
    ("hat_2d_f64", "float64", [5, 4], [33, 17], [0, 1], "hat"),
    ("bio_3d_f64", "float64", [3, 4, 2], [8, 17, 4], [0, 0, 1], "biorthogonal"),
    ("fw_2d_f32", "float32", [4, 4], [16, 16], [1, 0], "fullweighting"),
    ("bio_1d_f64", "float64", [7], [129], [2], "biorthogonal"),
]


def make_regression(oracle):
    manifest = []
    for k, (name, dtype, level, extents, lmin, wavelet) in enumerate(REGRESSION):
        rng = np.random.default_rng(1000 + k)
        a = rng.uniform(-1.0, 1.0, size=extents).astype(dtype)
        # close the periodic images: index 2^l equals index 0
        for d, (l, e) in enumerate(zip(level, extents)):
            if e == (1 << l) + 1:
                idx_img = [slice(None)] * len(extents)
                idx_src = [slice(None)] * len(extents)
                idx_img[d] = 1 << l
                idx_src[d] = 0
                a[tuple(idx_img)] = a[tuple(idx_src)]
        a = np.ascontiguousarray(a)
        fwd = a.copy()
        oracle.hierarchize(fwd, level, lmin, level, wavelet)
        back = fwd.copy()
        oracle.dehierarchize(back, level, lmin, level, wavelet)
        for tag, arr in (("input", a), ("hierarchized", fwd), ("dehierarchized", back)):
            arr.tofile(os.path.join(HERE, f"regression_{name}_{tag}.raw"))
        manifest.append({
            "name": name, "dtype": dtype, "level": level, "extents": extents,
            "minimum_level": lmin, "wavelet": wavelet,
        })
    with open(os.path.join(HERE, "regression.json"), "w") as f:
        json.dump(manifest, f, indent=1)


def main():
    with open(os.path.join(HERE, "plan_multipass.json"), "w") as f:
        json.dump(plan_multipass_cases(), f, indent=0)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bridge

    make_regression(oracle_bridge)


if __name__ == "__main__":
    main()
