"""Small instance generators and the half-integral enumeration oracle.

The PRAQ enumeration follows qubo/praq-cuda/tests/test_solver.cu:76-98: the
relaxation's vertices are half-integral (roof duality), so on a tiny instance
its optimum is min F(x) over {0, 1/2, 1}^n, F recomputed from the dense
definition (qubo/praq-jax/praq/qubo_problem.py:95-111).
"""
from __future__ import annotations

import itertools

import numpy as np


def random_qubo(seed: int, n: int, density: float, magnitude: int = 3):
    """Integer entries in +-[1, magnitude], off-diagonal kept with probability
    `density`; (n, diag, rows, cols, vals)."""
    rng = np.random.default_rng(seed)
    diag = rng.integers(1, magnitude + 1, n) * rng.choice([-1, 1], n)
    pairs = [(i, j) for i in range(n) for j in range(i + 1, n) if rng.random() < density]
    rows = np.array([p[0] for p in pairs], dtype=np.int64)
    cols = np.array([p[1] for p in pairs], dtype=np.int64)
    vals = (rng.integers(1, magnitude + 1, len(pairs)) * rng.choice([-1, 1], len(pairs)))
    return n, diag.astype(float), rows, cols, vals.astype(float)


def relaxation_value(n, diag, rows, cols, vals, x) -> float:
    x = np.asarray(x, dtype=float)
    w = 2.0 * np.asarray(vals)
    xi, xj = x[rows], x[cols]
    edge = np.where(w > 0, w * np.maximum(xi + xj - 1.0, 0.0), w * np.minimum(xi, xj))
    return float(np.asarray(diag) @ x + edge.sum())


def half_integral_optimum(n, diag, rows, cols, vals) -> float:
    best = np.inf
    for x in itertools.product((0.0, 0.5, 1.0), repeat=n):
        best = min(best, relaxation_value(n, diag, rows, cols, vals, x))
    return best


def small_jrp(seed: int, n: int, t: int):
    """Integer-cost JRP instance with the structure of pria-jax's generator."""
    rng = np.random.default_rng(seed)
    mult = rng.uniform(0, 1, n)
    s = np.arange(t)[:, None]
    tt = np.arange(t)[None, :]
    h = np.round(np.triu(mult[:, None, None] * (tt - s)[None]
                         + rng.uniform(0, 1, (n, t, t)) * mult[:, None, None], 1))
    k = np.maximum(1, np.round(rng.uniform(1, max(t, 1.01), n)))
    k0 = float(np.round(rng.uniform(1, max(t, 1.01)))) if n > 1 else 0.0
    return k.astype(float), k0, h


def brute_force_optima(n, diag, rows, cols, vals):
    """Every binary minimiser of sum d_i x_i + sum 2 W_ij x_i x_j."""
    xs = ((np.arange(2 ** n)[:, None] >> np.arange(n)) & 1).astype(float)
    f = xs @ np.asarray(diag) + (2.0 * np.asarray(vals) * xs[:, rows] * xs[:, cols]).sum(axis=1)
    return xs[np.abs(f - f.min()) < 1e-9]


def check_answer_properties(inst, x, y):
    """qubo/praq-jax/tests/test_solver.py:71-134: the primal is optimal for the
    relaxation (half-integral enumeration); the coordinates persistency fixes
    (tolerance 1e-5) agree with some binary optimum (brute force); covering
    duals have equal slots in [-w, 0], packing duals are nonnegative and sum
    to w (instance units, w = 2 |W_ij|)."""
    from reformulation_amd import praq
    n, diag, rows, cols, vals = inst
    opt = half_integral_optimum(*inst)
    assert abs(relaxation_value(*inst, x) - opt) <= 1e-6 * (1 + abs(opt))
    # the package's post-solve functions (praq-jax relaxation_bounds and
    # persistent_variables, solver.py:348-362, 592-611): the upper bound is
    # F(x), the lower bound a weak-duality bound that the solve has closed
    problem = praq.QuboProblem.from_sparse(n, diag, rows, cols, vals)
    lower, upper = praq.relaxation_bounds(problem, x, y)
    assert abs(upper - relaxation_value(*inst, x)) <= 1e-9 * (1 + abs(opt))
    assert lower <= opt + 1e-9 * (1 + abs(opt)) and opt - lower <= 1e-6 * (1 + abs(opt))
    fixed, rounded = praq.persistent_variables(x, tolerance=1e-5)
    x = np.asarray(x, dtype=float)
    assert np.array_equal(fixed, np.abs(x - np.round(x)) <= 1e-5)
    optima = brute_force_optima(n, diag, rows, cols, vals)
    assert (optima[:, fixed] == rounded[fixed]).all(axis=1).any()
    y = np.asarray(y, dtype=float).reshape(-1, 2)
    w = 2.0 * np.abs(np.asarray(vals))
    cover = np.asarray(vals) > 0
    yc, yp = y[cover], y[~cover]
    big = 1.0 + w.max(initial=0.0)
    assert np.abs(yc[:, 0] - yc[:, 1]).max(initial=0.0) <= 1e-9 * big
    assert (yc[:, 0] <= 1e-9).all() and (yc[:, 0] >= -w[cover] - 1e-9).all()
    assert np.abs(yp.sum(axis=1) - w[~cover]).max(initial=0.0) <= 1e-9 * big
    assert (yp >= -1e-9).all()
    return int(fixed.sum())
