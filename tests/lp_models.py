"""Independent references for the LP values: explicit LPs solved by HiGHS
(scipy.optimize.linprog) and a max-flow roof dual.  Test infrastructure.

PRAQ, the standard linearization of min x^T W x (the model of
qubo/baselines/model.py:1-22, built here independently):
    min <d, x> + sum_e w_e z_e,  w_e = 2 W_ij,
    x_i + x_j - z_e <= 1 (w_e > 0),   z_e <= x_i, z_e <= x_j (w_e < 0),
    0 <= x, z <= 1.
PRAQ max-flow, the Hammer-Hansen-Simeone implication network (derivation
summarised at qubo/baselines/maxflow_backend.py:19-60): write the objective as
a posiform -- w x_i x_j = |w| x_i (1 - x_j) - |w| x_i for w < 0, a x_i =
a + |a| (1 - x_i) for a < 0 -- and send each term a u v to arcs (not u) -> v
and (not v) -> u of capacity a / 2; the roof dual is constant + min cut.
Capacities are doubled to integers (exact for integer W) and the flow run on
scipy's int32 maximum_flow.
PRIA, the JRP LP of jrp/pria-cuda/include/pria/pria.h:4-14 with y_is, z_s:
    min K0 sum z_s + sum K_i y_is + sum H x,  sum_{s<=t} x_ist = 1,
    x_ist <= y_is, x_ist <= z_s, all >= 0.
"""
from __future__ import annotations

import numpy as np
from scipy.optimize import linprog
from scipy.sparse import coo_array, csr_array
from scipy.sparse.csgraph import maximum_flow


def praq_lp_value(inst) -> float:
    n = inst.n
    w = 2.0 * np.asarray(inst.vals, dtype=float)
    keep = w != 0.0
    i, j, w = np.asarray(inst.rows)[keep], np.asarray(inst.cols)[keep], w[keep]
    m = w.size
    cost = np.concatenate([np.asarray(inst.diag, dtype=float), w])
    cov = np.flatnonzero(w > 0)
    pac = np.flatnonzero(w < 0)
    r, c, v, rhs = [], [], [], []
    row = 0
    for e in cov:
        r += [row, row, row]
        c += [i[e], j[e], n + e]
        v += [1.0, 1.0, -1.0]
        rhs.append(1.0)
        row += 1
    for e in pac:
        for end in (i[e], j[e]):
            r += [row, row]
            c += [n + e, end]
            v += [1.0, -1.0]
            rhs.append(0.0)
            row += 1
    a = coo_array((v, (r, c)), shape=(row, n + m)).tocsr() if row else None
    res = linprog(cost, A_ub=a, b_ub=np.array(rhs) if row else None,
                  bounds=[(0.0, 1.0)] * (n + m), method="highs",
                  options=dict(dual_feasibility_tolerance=1e-10,
                               primal_feasibility_tolerance=1e-10))
    assert res.status == 0, res.message
    return float(res.fun)


def praq_maxflow_value(inst) -> float:
    """constant + min cut / 2 with doubled (integer) capacities; needs
    integer diag and 2 W_ij."""
    n = inst.n
    src, snk = 2 * n, 2 * n + 1
    lit = lambda k, neg: k + n if neg else k          # noqa: E731
    arcs: dict[tuple[int, int], int] = {}
    constant = 0.0

    def add(p, q, cap):
        arcs[(p, q)] = arcs.get((p, q), 0) + cap

    def neg(node):
        if node == snk:
            return src
        if node == src:
            return snk
        return node + n if node < n else node - n

    def term(a, u, v):
        # term a u v -> arcs (not u) -> v and (not v) -> u, capacity a / 2
        # each; stored doubled (a), so the min cut is twice the original
        add(neg(u), v, a)
        add(neg(v), u, a)

    linear = np.array(inst.diag, dtype=float).copy()
    for a_, b_, wv in zip(np.asarray(inst.rows), np.asarray(inst.cols), 2.0 * np.asarray(inst.vals)):
        wi = int(round(wv))
        if wi > 0:
            term(wi, lit(a_, False), lit(b_, False))
        elif wi < 0:
            term(-wi, lit(a_, False), lit(b_, True))
            linear[a_] += wi
    for k, a in enumerate(linear):
        ai = int(round(a))
        if ai > 0:
            term(ai, lit(k, False), snk)
        elif ai < 0:
            constant += ai
            term(-ai, lit(k, True), snk)
    keys = list(arcs.keys())
    rows = np.array([p for p, _ in keys], dtype=np.int32)
    cols = np.array([q for _, q in keys], dtype=np.int32)
    caps = np.array([arcs[k] for k in keys], dtype=np.int32)
    graph = csr_array((caps, (rows, cols)), shape=(2 * n + 2, 2 * n + 2))
    flow = maximum_flow(graph, src, snk).flow_value
    # each term's two arcs carry a/2 each in the undoubled network; doubled
    # caps give twice the cut, and the constant is in original units
    return constant + flow / 2.0


def pria_lp_value(inst) -> float:
    k = np.asarray(inst.order_costs, dtype=float)
    k0 = float(inst.joint_order_cost)
    h = np.asarray(inst.holding_costs, dtype=float)
    n, t = h.shape[0], h.shape[1]
    xs = [(i, s, tt) for i in range(n) for s in range(t) for tt in range(s, t)]
    nx = len(xs)
    ny = n * t
    nz = t
    xi = {key: q for q, key in enumerate(xs)}
    cost = np.concatenate([[h[i, s, tt] for i, s, tt in xs], np.repeat(k, t), np.full(t, k0)])
    er, ec, ev, eb = [], [], [], []
    row = 0
    for i in range(n):
        for tt in range(t):
            for s in range(tt + 1):
                er.append(row); ec.append(xi[(i, s, tt)]); ev.append(1.0)
            eb.append(1.0)
            row += 1
    a_eq = coo_array((ev, (er, ec)), shape=(row, nx + ny + nz)).tocsr()
    ur, uc, uv = [], [], []
    row = 0
    for q, (i, s, tt) in enumerate(xs):
        ur += [row, row]; uc += [q, nx + i * t + s]; uv += [1.0, -1.0]
        row += 1
        ur += [row, row]; uc += [q, nx + ny + s]; uv += [1.0, -1.0]
        row += 1
    a_ub = coo_array((uv, (ur, uc)), shape=(row, nx + ny + nz)).tocsr()
    res = linprog(cost, A_ub=a_ub, b_ub=np.zeros(row), A_eq=a_eq, b_eq=np.array(eb),
                  bounds=[(0.0, None)] * (nx + ny + nz), method="highs",
                  options=dict(dual_feasibility_tolerance=1e-10,
                               primal_feasibility_tolerance=1e-10))
    assert res.status == 0, res.message
    return float(res.fun)
