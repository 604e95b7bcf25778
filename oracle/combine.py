"""CPU oracle of the combination-technique combine -- TEST INFRASTRUCTURE ONLY.

Restates, with numpy, the steps that follow the batched hierarchize in the
reference's combination-technique workflow, so that the HIP combine can be
checked bit for bit.  Only the tests, the smoke check and the benchmark's CPU
baseline may import it; the product path never does.

Restated reference code (file:line in freifrauvonbleifrei/paliwa):
  * iterate_hierarchical_subspaces        paliwa/paliwa_utils.hpp:9-25
      every subspace level k with 0 <= k_d <= l_d of a component grid l
  * subspace counting and the count > 1 rule
                                          test/test_paliwa_combination_technique.cpp:176-193, 217-233
      only subspaces contained in more than one component grid are kept
  * strided_hierarchical_domain_from_level paliwa/paliwa_domains.hpp:51-83
      the points of subspace k along d, in full-grid index space at maximum
      level M: k = 0 -> {0}; k >= 1 -> (2m+1) 2^(M-k), m in [0, 2^(k-1))
  * strided_domain_from_level             paliwa/paliwa_domains.hpp:33-49
      a component grid at level l stores full-grid index i at i / 2^(M-l)
  * the collect loop                      test/test_paliwa_combination_technique.cpp:254-290
      per grid in index order: subspace(i) += coefficient * grid(i)
      (the product rounded, then the sum: no fused multiply-add)
  * the scatter onto the full grid        test/test_paliwa_combination_technique.cpp:292-313
      a zero-filled full grid at the maximum level, full(i) += subspace(i)

Subspaces are ordered lexicographically by level vector (the reference's
std::map order) and stored one after another, each row-major over its
2^(k_d - 1) (or 1) points per axis.  The reference assigns storage in the
iteration order of a Kokkos::UnorderedMap, which no caller can observe; only
the values at each point are compared.

Pinned by tests/test_combine_oracle.py: the reference test's grid and sparse
sizes (test/test_paliwa_combination_technique.cpp:137-145, 226-232), its
mean-value check of the combined, dehierarchized full grid (:333-340), and the
hierarchical-surplus identity (with the hat basis and minimum level 0 every
component grid holds the full grid's surplus at each of its points, so the
combined subspace equals the full grid's surplus times the fold of the
coefficients), plus a brute-force point-by-point restatement on tiny sets.
"""
from __future__ import annotations

import itertools

import numpy as np


def subspace_counts(levels):
    """{k: number of component grids containing subspace k}."""
    counts = {}
    for l in levels:
        for k in itertools.product(*[range(x + 1) for x in l]):
            counts[k] = counts.get(k, 0) + 1
    return counts


def counted_subspaces(levels, min_count=2):
    """Subspaces contained in at least `min_count` grids, lexicographic."""
    return sorted(k for k, c in subspace_counts(levels).items() if c >= min_count)


def subspace_shape(k):
    return tuple(1 if x == 0 else 1 << (x - 1) for x in k)


def subspace_offsets(subspaces):
    sizes = [int(np.prod(subspace_shape(k))) for k in subspaces]
    return np.concatenate([[0], np.cumsum(sizes, dtype=np.int64)]).astype(np.int64)


def _component_slices(k, l):
    """Index slices of subspace k inside a component grid stored at level l."""
    out = []
    for kd, ld in zip(k, l):
        if kd == 0:
            out.append(slice(0, 1))
        else:
            out.append(slice(1 << (ld - kd), 1 << ld, 1 << (ld - kd + 1)))
    return tuple(out)


def _full_slices(k, maximum_level):
    out = []
    for kd, md in zip(k, maximum_level):
        if kd == 0:
            out.append(slice(0, 1))
        else:
            out.append(slice(1 << (md - kd), 1 << md, 1 << (md - kd + 1)))
    return tuple(out)


def gather(grids, levels, coefficients, min_count=2):
    """Sparse-grid storage: for each counted subspace, the fold over the
    grids containing it (index order) of  s += coefficient * grid.
    Returns (sparse, subspaces, offsets)."""
    subspaces = counted_subspaces(levels, min_count)
    offsets = subspace_offsets(subspaces)
    dtype = grids[0].dtype
    sparse = np.zeros(int(offsets[-1]), dtype=dtype)
    for si, k in enumerate(subspaces):
        s = np.zeros(subspace_shape(k), dtype=dtype)
        for g, l, c in zip(grids, levels, coefficients):
            if any(kd > ld for kd, ld in zip(k, l)):
                continue
            v = g[_component_slices(k, l)]
            if dtype == np.float64:
                s = s + np.float64(c) * v
            else:
                # float grid, double coefficient: the product and the sum in
                # double, the sum stored back to float
                s = (s.astype(np.float64) + np.float64(c) * v.astype(np.float64)).astype(dtype)
        sparse[offsets[si]:offsets[si + 1]] = s.ravel()
    return sparse, subspaces, offsets


def scatter(sparse, subspaces, offsets, maximum_level, shape=None):
    """The full grid at the maximum level: zero, plus each counted subspace at
    its points.  `shape`: 2^M or 2^M+1 per axis (the image plane, if any,
    repeats index 0)."""
    M = list(maximum_level)
    lattice = tuple(1 << m for m in M)
    shape = lattice if shape is None else tuple(shape)
    full = np.zeros(lattice, dtype=sparse.dtype)
    for si, k in enumerate(subspaces):
        block = sparse[offsets[si]:offsets[si + 1]].reshape(subspace_shape(k))
        full[_full_slices(k, M)] = full[_full_slices(k, M)] + block
    if shape == lattice:
        return full
    out = np.zeros(shape, dtype=sparse.dtype)
    for corner in itertools.product(*[(False, True) if s == n + 1 else (False,)
                                      for s, n in zip(shape, lattice)]):
        dst = tuple(slice(n, n + 1) if c else slice(0, n) for c, n in zip(corner, lattice))
        src = tuple(slice(0, 1) if c else slice(0, n) for c, n in zip(corner, lattice))
        out[dst] = full[src]
    return out


def gather_pointwise(grids, levels, coefficients, maximum_level, min_count=2):
    """Brute force of gather for tiny sets: per full-grid point, its subspace
    from the point's own coordinates, then the fold over the grids that hold
    the point.  Returns {full-grid index tuple: value}."""
    counts = subspace_counts(levels)
    out = {}
    for i in itertools.product(*[range(1 << m) for m in maximum_level]):
        k = []
        for x, m in zip(i, maximum_level):
            if x == 0:
                k.append(0)
            else:
                tz = (x & -x).bit_length() - 1
                k.append(m - tz)
        k = tuple(k)
        if counts.get(k, 0) < min_count:
            continue
        s = 0.0
        for g, l, c in zip(grids, levels, coefficients):
            if any(kd > ld for kd, ld in zip(k, l)):
                continue
            j = tuple(x >> (m - ld) for x, m, ld in zip(i, maximum_level, l))
            s = s + float(c) * float(g[j])
        out[i] = s
    return out
