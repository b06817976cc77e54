"""Seeded synthetic workloads: the instance distributions the measurements
are quoted on.  The reference has no generator for them, so each is stated
here and drawn with numpy's default_rng(seed).

  praq_er           Erdos-Renyi coupling graph with m distinct pairs i < j,
                    q_ij uniform over {-100..100} \\ {0}, c_i uniform over
                    {-100..100}; W_ii = c_i and W_ij = q_ij / 2, since x^T W x
                    puts 2 W_ij on each edge (n = 2000, m = 8000: plain PRAQ)
  praq_chung_lu     power-law expected degrees w_i ~ (i + i0)^(-1/(beta-1))
                    scaled to a mean degree, q_ij, c_i ~ N(0, 1)
                    (n = 1e6, beta 2.5, mean degree 32: the precision study)
  pria_c1           N items, T periods, d_it uniform over {1..50} with 20% set
                    to 0, h_i ~ U[0.1, 1], K_i ~ U[10, 100], K0 ~ U[100, 500],
                    H_ist = h_i d_it (t - s) (N = 50, T = 200: plain PRIA)
  mix               the batched mixture: 70% PRAQ with n log-uniform in
                    [1e3, 1e5], mean degree U[4, 32], integer weights as
                    praq_er; 30% PRIA with N uniform in {10..100}, T uniform
                    in {50..300}, costs as pria_c1; instance k has seed k
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .praq import QuboProblem
from .pria import JrpProblem


@dataclass(frozen=True)
class QuboInstance:
    """The reference harness's QUBO `Instance` (qubo/harness/instances.py:18-68)."""

    name: str
    family: str
    n: int
    diag: np.ndarray
    rows: np.ndarray
    cols: np.ndarray
    vals: np.ndarray

    @property
    def edges(self) -> int:
        return int(self.vals.size)

    def problem(self) -> QuboProblem:
        return QuboProblem.from_sparse(self.n, self.diag, self.rows, self.cols, self.vals)


@dataclass(frozen=True)
class JrpInstance:
    """The reference harness's JRP `Instance` (jrp/harness/instances.py:27-78)."""

    name: str
    family: str
    order_costs: np.ndarray
    joint_order_cost: float
    holding_costs: np.ndarray

    @property
    def items(self) -> int:
        return int(self.order_costs.size)

    @property
    def periods(self) -> int:
        return int(self.holding_costs.shape[-1])

    def problem(self) -> JrpProblem:
        return JrpProblem.new(self.order_costs, self.joint_order_cost, self.holding_costs)


def _distinct_pairs(rng, n: int, m: int) -> tuple[np.ndarray, np.ndarray]:
    """m distinct unordered pairs i < j, uniform, sorted by (i, j)."""
    m = min(m, n * (n - 1) // 2)
    chosen = np.empty(0, dtype=np.int64)
    while chosen.size < m:
        k = int((m - chosen.size) * 1.2) + 16
        i = rng.integers(0, n, k)
        j = rng.integers(0, n, k)
        keep = i != j
        lo, hi = np.minimum(i, j)[keep], np.maximum(i, j)[keep]
        chosen = np.unique(np.concatenate([chosen, lo.astype(np.int64) * n + hi]))
    if chosen.size > m:
        chosen = np.sort(rng.choice(chosen, size=m, replace=False))
    return chosen // n, chosen % n


def _int_weights(rng, m: int) -> np.ndarray:
    q = rng.integers(1, 101, m) * np.where(rng.random(m) < 0.5, -1, 1)
    return q.astype(np.float64)


def praq_er(seed: int, n: int = 2000, m: int = 8000) -> QuboInstance:
    rng = np.random.default_rng(seed)
    rows, cols = _distinct_pairs(rng, n, m)
    q = _int_weights(rng, rows.size)
    c = rng.integers(-100, 101, n).astype(np.float64)
    return QuboInstance(f"er-n{n}-m{m}-s{seed}", "er", n, c, rows, cols, 0.5 * q)


def praq_chung_lu(seed: int = 0, n: int = 1_000_000, beta: float = 2.5,
                  mean_degree: float = 32.0) -> QuboInstance:
    """Chung-Lu graph: pairs drawn with probability proportional to w_i w_j,
    by sampling both ends from the degree distribution; duplicates and
    self-loops dropped."""
    rng = np.random.default_rng(seed)
    i0 = 1.0
    w = (np.arange(n) + i0) ** (-1.0 / (beta - 1.0))
    w *= mean_degree * n / w.sum()
    target = int(round(mean_degree * n / 2))
    p = w / w.sum()
    cdf = np.cumsum(p)
    chosen = np.empty(0, dtype=np.int64)
    draws = 0
    while chosen.size < target and draws < 8:
        k = int((target - chosen.size) * 1.3) + 16
        a = np.minimum(np.searchsorted(cdf, rng.random(k)), n - 1)
        b = np.minimum(np.searchsorted(cdf, rng.random(k)), n - 1)
        keep = a != b
        lo, hi = np.minimum(a, b)[keep], np.maximum(a, b)[keep]
        chosen = np.unique(np.concatenate([chosen, lo.astype(np.int64) * n + hi]))
        draws += 1
    if chosen.size > target:
        chosen = np.sort(rng.choice(chosen, size=target, replace=False))
    rows, cols = chosen // n, chosen % n
    vals = rng.standard_normal(rows.size)
    diag = rng.standard_normal(n)
    return QuboInstance(f"chunglu-n{n}-s{seed}", "chung-lu", n, diag, rows, cols, vals)


def pria_c1(seed: int, items: int = 50, periods: int = 200) -> JrpInstance:
    rng = np.random.default_rng(seed)
    d = rng.integers(1, 51, (items, periods)).astype(np.float64)
    d[rng.random((items, periods)) < 0.2] = 0.0
    h = rng.uniform(0.1, 1.0, items)
    k = rng.uniform(10.0, 100.0, items)
    k0 = float(rng.uniform(100.0, 500.0))
    lag = np.maximum(np.arange(periods)[None, :] - np.arange(periods)[:, None], 0)
    H = h[:, None, None] * d[:, None, :] * lag[None, :, :]
    return JrpInstance(f"jrp-n{items}-t{periods}-s{seed}", "jrp", k, k0, H)


def mix_member(seed: int):
    """The mixture's instance `seed`: ("praq", QuboInstance) or ("pria", JrpInstance)."""
    rng = np.random.default_rng(10_000_019 + seed)
    if rng.random() < 0.7:
        n = int(round(math.exp(rng.uniform(math.log(1e3), math.log(1e5)))))
        degree = rng.uniform(4.0, 32.0)
        m = int(round(n * degree / 2))
        inst = praq_er(seed, n, m)
        return "praq", inst
    items = int(rng.integers(10, 101))
    periods = int(rng.integers(50, 301))
    return "pria", pria_c1(seed, items, periods)


def mix(count: int = 1024, start: int = 0):
    return [mix_member(k) for k in range(start, start + count)]
