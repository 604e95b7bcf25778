"""Combination-technique sets of the reference's integration test
(test/test_paliwa_combination_technique.cpp:88-124) and its field
(:41-73), shared by the oracle and GPU combine tests."""
from __future__ import annotations

import math

import numpy as np

# dimension -> (maximum_level, minimum_level, component levels, coefficients)
REFERENCE_SETS = {
    2: ((10, 11), (2, 3),
        [(2, 11), (3, 10), (4, 9), (5, 8), (6, 7), (7, 6), (8, 5), (9, 4), (10, 3),
         (2, 10), (3, 9), (4, 8), (5, 7), (6, 6), (7, 5), (8, 4), (9, 3)],
        [1] * 9 + [-1] * 8),
    3: ((5, 6, 7), (4, 5, 6),
        [(4, 6, 7), (5, 5, 7), (5, 6, 6), (4, 5, 6)],
        [1, 1, 1, -2]),
    4: ((3, 4, 5, 6), (1, 2, 3, 4),
        [(2, 2, 4, 4), (3, 2, 3, 4), (1, 4, 3, 4), (1, 2, 5, 4), (2, 3, 3, 4),
         (1, 3, 4, 4), (2, 2, 3, 5), (1, 3, 3, 5), (1, 2, 3, 6), (1, 2, 4, 5),
         (1, 2, 4, 4), (2, 2, 3, 4), (1, 3, 3, 4), (1, 2, 3, 5), (1, 2, 3, 4)],
        [1] * 10 + [-3] * 4 + [3]),
}

# test/test_paliwa_combination_technique.cpp:137-145 and 226-232
REFERENCE_GRID_POINTS = {2: 106496, 3: 425984, 4: 50176}
REFERENCE_SPARSE_POINTS = {2: 18432, 3: 131072, 4: 5120}


def sine_product_grid(level, dtype=np.float64):
    """prod_d sin(pi x_d), x_d = j / 2^l_d on the component grid's lattice
    (the full-grid coordinate i / 2^M of the point it stores)."""
    axes = [np.sin(math.pi * np.arange(1 << l) / float(1 << l)) for l in level]
    g = np.ones([1 << l for l in level])
    for d, a in enumerate(axes):
        shape = [1] * len(level)
        shape[d] = a.size
        g = g * a.reshape(shape)
    return np.ascontiguousarray(g.astype(dtype))


def standard_set(dim, n, minimum=1):
    """The classical combination technique: levels with l_i >= minimum and
    |l|_1 = n - q, q = 0..dim-1, coefficient (-1)^q C(dim-1, q)."""
    import itertools

    levels, coefs = [], []
    for q in range(dim):
        total = n - q
        for l in itertools.product(range(minimum, total + 1), repeat=dim):
            if sum(l) == total:
                levels.append(tuple(l))
                coefs.append((-1) ** q * math.comb(dim - 1, q))
    return levels, coefs
