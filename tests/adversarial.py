"""Designed JRP instances with known LP values, for the parity tests.  Test
infrastructure only.

The reference builds these by KKT inverse design
(jrp/pria-jax/pria/adversarial.py:45-90 tables, :120-260 constructions) and
its tests assert their values (jrp/pria-jax/tests/test_adversarial.py:24-90).
The 0/1 gadget tables and bump sets below are that data; the construction is
restated so the same integer instances come out and the known answers carry
over:

* a gadget Z (p x p, 0/1) has a fractional vertex mu = Z^-1 1 in (0, 1)^p
  with denominator |det Z|;
* the single-item design places Z on periods: anchor period 1 serves periods
  1..p+1, "star" periods 2..p+1 serve the gadget periods p+2..2p+1 where Z
  has a one, each gadget period also serves itself; holding costs are
  (K + 1) - multiplier + slack, so that the designed point is optimal;
* the tie design puts two gadgets on shared gadget periods, and every point
  of the face Z_a mu_a + Z_b mu_b = 1 then has value (K + 1) T; integer
  bumps on support entries split the tie;
* a single item is embedded as item 1 of N = 2 beside a decoy item (holding 0
  on the diagonal, 50 off it, K = 1) with K0 = 1: the decoy orders every
  period, so z = 1 throughout and the full instance's LP value is item 1's
  plus 2 T.
"""
from __future__ import annotations

import numpy as np

# the shipped gadgets and partners (adversarial.py:52-77)
GADGETS = {
    4: ((0, 1, 1, 1), (1, 0, 1, 1), (1, 1, 0, 1), (1, 1, 1, 0)),
    6: ((0, 1, 1, 1, 0, 1), (1, 0, 1, 0, 1, 0), (1, 0, 0, 1, 0, 1),
        (0, 0, 0, 0, 1, 1), (1, 1, 0, 1, 1, 0), (1, 1, 0, 0, 0, 1)),
    7: ((0, 1, 0, 0, 1, 1, 1), (1, 0, 1, 0, 0, 1, 1), (1, 0, 0, 1, 1, 1, 0),
        (0, 0, 1, 1, 1, 1, 0), (1, 1, 1, 0, 0, 0, 0), (1, 0, 1, 0, 1, 0, 1),
        (0, 0, 0, 1, 0, 0, 1)),
    8: ((1, 1, 1, 0, 0, 0, 0, 0), (1, 1, 0, 1, 1, 0, 0, 1), (0, 1, 1, 0, 1, 1, 0, 1),
        (0, 1, 1, 0, 1, 0, 1, 1), (0, 0, 1, 1, 0, 1, 1, 1), (1, 0, 1, 1, 1, 0, 0, 0),
        (1, 0, 0, 0, 1, 1, 1, 0), (0, 1, 0, 1, 1, 1, 1, 0)),
}
PARTNER_8 = ((1, 0, 0, 0, 1, 1, 0, 1), (0, 1, 1, 0, 1, 0, 1, 0), (0, 0, 1, 1, 0, 1, 0, 0),
             (0, 1, 0, 1, 0, 0, 0, 1), (1, 0, 1, 0, 1, 0, 0, 1), (1, 1, 0, 1, 1, 0, 0, 0),
             (1, 0, 0, 1, 1, 0, 1, 1), (1, 1, 1, 0, 0, 0, 1, 0))
PARTNER_4 = ((1, 0, 0, 1), (0, 1, 0, 1), (0, 1, 1, 0), (1, 1, 0, 0))
# the slowest known bump set of the p = 8 pair, exact margin 1/145 (adversarial.py:88-90)
HARD_BUMPS = ((2, 18), (6, 24), (7, 19), (15, 24))


def gadget(p: int) -> np.ndarray:
    return np.array(GADGETS[p], dtype=int)


def _shares(total: int, parts: int) -> list[int]:
    """total split into `parts` integers, the first ones one larger."""
    q, r = divmod(total, parts)
    return [q + (j < r) for j in range(parts)]


def _holding(T: int, K: int, multipliers: dict, support: set, stars: set,
             gadget_periods: set) -> np.ndarray:
    """One-based (T + 1) x (T + 1) holding costs from the design's multipliers."""
    H = np.zeros((T + 1, T + 1), dtype=int)
    for t in range(1, T + 1):
        for s in range(1, t + 1):
            if (s, t) in support:
                slack = 0
            elif s in stars and t in gadget_periods:
                slack = 1
            else:
                slack = 5
            H[s, t] = (K + 1) - multipliers.get((s, t), 0) + slack
    return H


def _place(Z, star_periods, gadget_periods, K, multipliers, support):
    for col, s in enumerate(star_periods):
        served = [t for row, t in enumerate(gadget_periods) if Z[row][col]]
        for t, m in zip(served, _shares(K, len(served))):
            multipliers[(s, t)] = m
            support.add((s, t))


def single_item_design(Z: np.ndarray):
    """(T, K, H one-based) of the fractional-optimum design, T = 2p + 1."""
    p = len(Z)
    T, K = 2 * p + 1, p + 1
    stars = list(range(2, p + 2))
    gperiods = list(range(p + 2, 2 * p + 2))
    mult = {(1, t): 1 for t in range(1, p + 2)}
    support = {(1, t) for t in range(1, p + 2)}
    _place(Z, stars, gperiods, K, mult, support)
    for s in gperiods:
        mult[(s, s)] = K
    return T, K, _holding(T, K, mult, support, set(stars), set(gperiods))


def tie_design(Za: np.ndarray, Zb: np.ndarray, bumps=()):
    """(T, K, H one-based) of the tie design, T = 3p + 1, bumps applied."""
    p = len(Za)
    T, K = 3 * p + 1, p + 1
    stars_a = list(range(2, p + 2))
    stars_b = list(range(p + 2, 2 * p + 2))
    gperiods = list(range(2 * p + 2, 3 * p + 2))
    mult = {(1, t): 1 for t in range(1, K + 1)}
    support = {(1, t) for t in range(1, 2 * p + 2)}
    _place(Za, stars_a, gperiods, K, mult, support)
    _place(Zb, stars_b, gperiods, K, mult, support)
    for s in gperiods:
        mult[(s, s)] = K
    H = _holding(T, K, mult, support, set(stars_a + stars_b), set(gperiods))
    stars = stars_a + stars_b
    for j, t in bumps:
        assert (stars[j], t) in support
        H[stars[j], t] += 1
    return T, K, H


def designed_value(T: int, K: int, H: np.ndarray, Z: np.ndarray) -> float:
    """Item 1's LP value at the designed vertex (anchor mass, the gadget
    vertex's holding and order costs)."""
    p = len(Z)
    mu = np.linalg.solve(Z, np.ones(p))
    v = float(sum(H[1, t] for t in range(1, p + 2)) + K)
    for col, s in enumerate(range(2, p + 2)):
        for row, t in enumerate(range(p + 2, 2 * p + 2)):
            if Z[row, col]:
                v += H[s, t] * mu[col]
        v += K * mu[col]
    return v


def embed(holding: np.ndarray, K: int, decoy: float = 50.0):
    """(order_costs, joint_cost, holding_costs) of the N = 2 instance."""
    T = holding.shape[-1]
    d = np.full((T, T), decoy)
    np.fill_diagonal(d, 0.0)
    h = np.stack([np.asarray(holding, dtype=float), np.triu(d)])
    return np.array([float(K), 1.0]), 1.0, h


def fractional_optimum(p: int):
    Z = gadget(p)
    T, K, H = single_item_design(Z)
    k, k0, h = embed(H[1:, 1:], K)
    return k, k0, h, dict(T=T, K=K, item=H[1:, 1:], value=designed_value(T, K, H, Z))


def tie_split(bumps=HARD_BUMPS, Za=None, Zb=None):
    Za = gadget(8) if Za is None else np.asarray(Za)
    Zb = np.array(PARTNER_8) if Zb is None else np.asarray(Zb)
    T, K, H = tie_design(Za, Zb, bumps)
    k, k0, h = embed(H[1:, 1:], K)
    return k, k0, h, dict(T=T, K=K, item=H[1:, 1:], tie=(K + 1) * T)


def single_item_lp(holding: np.ndarray, K: float):
    """min <H, x> + K sum_s mu_s over sum_{s<=t} x_st = 1, 0 <= x_st <= mu_s
    (HiGHS); returns (value, x over the pairs s <= t in (t, s) order)."""
    from scipy.optimize import linprog
    T = holding.shape[-1]
    pairs = [(s, t) for t in range(T) for s in range(t + 1)]
    nx = len(pairs)
    c = np.array([holding[s, t] for s, t in pairs] + [K] * T, dtype=float)
    a_eq = np.zeros((T, nx + T))
    a_ub = np.zeros((nx, nx + T))
    for q, (s, t) in enumerate(pairs):
        a_eq[t, q] = 1.0
        a_ub[q, q] = 1.0
        a_ub[q, nx + s] = -1.0
    r = linprog(c, A_ub=a_ub, b_ub=np.zeros(nx), A_eq=a_eq, b_eq=np.ones(T),
                bounds=[(0, None)] * (nx + T), method="highs")
    assert r.status == 0, r.message
    return float(r.fun), r.x[:nx]
