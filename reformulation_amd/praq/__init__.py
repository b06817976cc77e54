"""PRAQ: the QUBO roof-dual LP relaxation, solved on MI355X.

The entry points, options and return values mirror praq-jax
(qubo/praq-jax/praq/prelude.py:1-13, solver.py:43-611): `QuboProblem`,
`Parameters.with_auto_step`, `check_interval`, `solve_qubo_problem`,
`solve_qubo_problem_warm_start`, `relaxation_bounds`, `persistent_variables`.
Every solve runs the gfx950 kernels of libreformulation_hip.so through its
C ABI (`praq_solve_ex`); nothing here iterates on the host.

Differences from praq-jax, all additive:
  * instances may be sparse (`QuboProblem.from_sparse`), the form the
    reference harness stores; a dense `QuboProblem.new(matrix)` solves with
    praq-jax's dense-build degree (n - 1 on every node), so it takes
    praq-jax's steps, and returns the dual over all pairs i < j as praq-jax
    does;
  * `Parameters` carries the iterate precision (`use_float32`), the execution
    tier and a time limit.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, replace

import numpy as np

from .. import _native

INTERVAL_NODES = 50_000.0
INTERVAL_OFFSET = 0.0
INTERVAL_SCALE = 4.328085122666891   # 3 / ln 2
INTERVAL_POWER = 0.0
COUPLING_NORM = 1.0

TIER_AUTO, TIER_WORKGROUP, TIER_DEVICE = 0, 1, 2


@dataclass(frozen=True)
class Parameters:
    """praq-jax's `Parameters` (qubo/praq-jax/praq/solver.py:43-196)."""

    omega: float
    eta: float
    beta_sufficient: float
    beta_artificial: float
    beta_necessary: float
    theta: float
    theta_target: float
    theta_decay: float
    max_iter: int
    use_halpern: bool
    warm_steps: int
    op_steps: int
    tolerance: float
    interval_offset: float = INTERVAL_OFFSET
    interval_scale: float = INTERVAL_SCALE
    interval_power: float = INTERVAL_POWER
    # extensions
    use_float32: bool = False
    tier: int = TIER_AUTO
    time_limit: float = 0.0
    dense_degree: bool | None = None   # None: dense problems use n - 1

    @classmethod
    def new(cls, omega, eta, beta_sufficient, beta_artificial, beta_necessary, theta_initial,
            theta_target, theta_decay, max_iter, use_halpern, warm_steps, op_steps, tolerance,
            interval_offset=INTERVAL_OFFSET, interval_scale=INTERVAL_SCALE,
            interval_power=INTERVAL_POWER, **extensions) -> "Parameters":
        if max_iter is None:
            max_iter = -1
        op_steps = max(op_steps, 2) if op_steps > 0 else 0
        warm_steps = max(warm_steps, 0)
        return cls(float(omega), float(eta), float(beta_sufficient), float(beta_artificial),
                   float(beta_necessary), float(theta_initial), float(theta_target),
                   float(theta_decay), int(max_iter), bool(use_halpern), int(warm_steps),
                   int(op_steps), float(tolerance), float(interval_offset),
                   float(interval_scale), float(interval_power), **extensions)

    @classmethod
    def with_auto_step(cls, max_iter, tolerance, use_halpern=True, op_steps=16, omega=1.0,
                       beta_sufficient=float(np.exp(-1.0)), beta_necessary=0.5706,
                       beta_artificial=0.2468, theta_initial=0.81930, theta_target=0.42818,
                       theta_decay=0.90474, warm_steps=0, interval_offset=INTERVAL_OFFSET,
                       interval_scale=INTERVAL_SCALE, interval_power=INTERVAL_POWER,
                       **extensions) -> "Parameters":
        eta = 0.999 / COUPLING_NORM
        return cls.new(omega, eta, beta_sufficient, beta_artificial, beta_necessary,
                       theta_initial, theta_target, theta_decay, max_iter, use_halpern,
                       warm_steps, op_steps, tolerance, interval_offset, interval_scale,
                       interval_power, **extensions)

    @classmethod
    def stack(cls, parameters: list["Parameters"]) -> list["Parameters"]:
        """praq-jax stacks omega/theta for vmap; a batch here takes one
        Parameters per instance, so the list is returned as it is."""
        return list(parameters)


def check_interval(n: int, offset: float = INTERVAL_OFFSET, scale: float = INTERVAL_SCALE,
                   power: float = INTERVAL_POWER) -> int:
    """max(2, round(offset + scale B(ln(n / 50000), power))) with B the Box-Cox
    family (qubo/praq-jax/praq/solver.py:199-225); computed by the library."""
    p = _native.lib().praq_default_parameters()
    p.interval_offset, p.interval_scale, p.interval_power = offset, scale, power
    return int(_native.lib().praq_check_interval(int(n), ctypes.byref(p)))


class QuboProblem:
    """min_{x in {0,1}^n} x^T W x.

    Stored sparse as (diag, rows, cols, vals) with rows < cols, the layout of
    the reference harness's `Instance` (qubo/harness/instances.py:18-68).  A
    problem built with `new(matrix)` also remembers it was dense, which makes
    the solve use praq-jax's dense-build degree and return praq-jax's dual
    shape.
    """

    def __init__(self, n, diag, rows, cols, vals, dense: bool = False):
        self._n = int(n)
        self.diag = np.ascontiguousarray(diag, dtype=np.float64)
        self.rows = np.ascontiguousarray(rows, dtype=np.int64)
        self.cols = np.ascontiguousarray(cols, dtype=np.int64)
        self.vals = np.ascontiguousarray(vals, dtype=np.float64)
        self.dense = dense

    def n(self) -> int:
        return self._n

    def m(self) -> int:
        return int(self.vals.size)

    @classmethod
    def new(cls, matrix) -> "QuboProblem":
        """From a dense (n, n) W, symmetrized (qubo_problem.py:22-47)."""
        w = np.asarray(matrix, dtype=np.float64)
        w = 0.5 * (w + w.T)
        n = w.shape[0]
        r, c = np.triu_indices(n, k=1)
        keep = w[r, c] != 0.0
        return cls(n, np.diagonal(w).copy(), r[keep], c[keep], w[r, c][keep], dense=True)

    @classmethod
    def from_sparse(cls, n, diag, rows, cols, vals) -> "QuboProblem":
        return cls(n, diag, rows, cols, vals, dense=False)

    @classmethod
    def from_instance(cls, instance) -> "QuboProblem":
        """From a reference-harness `Instance` (or anything with n, diag,
        rows, cols, vals)."""
        return cls(instance.n, instance.diag, instance.rows, instance.cols, instance.vals)

    @classmethod
    def zero_instance(cls, n: int) -> "QuboProblem":
        return cls.new(np.zeros((n, n)))

    @classmethod
    def random_instance(cls, seed: int, n: int, density: float = 1.0,
                        magnitude: float = 3.0) -> "QuboProblem":
        """Integer entries uniform over +-[1, magnitude], off-diagonal kept
        with probability `density` -- the distribution of praq-jax's
        generator (qubo_problem.py:59-85), drawn with numpy."""
        rng = np.random.default_rng(seed)
        values = np.minimum(np.round(rng.uniform(1.0, magnitude + 1.0, (n, n))), magnitude)
        signs = np.where(rng.random((n, n)) < 0.5, 1.0, -1.0)
        keep = (rng.random((n, n)) < density) | np.eye(n, dtype=bool)
        m = np.where(keep, values * signs, 0.0)
        return cls.new(np.triu(m) + np.triu(m, 1).T)

    def to_dense(self) -> np.ndarray:
        w = np.zeros((self._n, self._n))
        w[np.arange(self._n), np.arange(self._n)] = self.diag
        w[self.rows, self.cols] = self.vals
        w[self.cols, self.rows] = self.vals
        return w

    def objective_scale(self) -> float:
        """||(d; 2|W_ij|)||_2, 1 for the zero instance (qubo_problem.py:113-124)."""
        s = math.sqrt(float(np.sum(self.diag ** 2) + np.sum((2.0 * self.vals) ** 2)))
        return s if s > 0.0 else 1.0

    def objective_value(self, x) -> float:
        x = np.asarray(x, dtype=np.float64)
        return float(self.diag @ x + 2.0 * np.sum(self.vals * x[self.rows] * x[self.cols]))

    def relaxation_value(self, x) -> float:
        """F(x) (qubo_problem.py:95-111)."""
        x = np.asarray(x, dtype=np.float64)
        xi, xj, w = x[self.rows], x[self.cols], 2.0 * self.vals
        edge = np.where(w > 0.0, w * np.maximum(xi + xj - 1.0, 0.0), w * np.minimum(xi, xj))
        return float(self.diag @ x + edge.sum())


def _native_params(L, params: Parameters, dense: bool):
    p = L.praq_default_parameters()
    p.tolerance = params.tolerance
    p.max_iter = params.max_iter if params.max_iter > 0 else -1
    p.use_float32 = int(params.use_float32)
    p.omega = params.omega
    p.beta_sufficient = params.beta_sufficient
    p.beta_necessary = params.beta_necessary
    p.beta_artificial = params.beta_artificial
    p.theta_initial = params.theta
    p.theta_target = params.theta_target
    p.theta_decay = params.theta_decay
    p.op_steps = params.op_steps
    p.interval_offset = params.interval_offset
    p.interval_scale = params.interval_scale
    p.interval_power = params.interval_power
    o = L.praq_default_options()
    o.use_halpern = int(params.use_halpern)
    o.warm_steps = params.warm_steps
    o.dense_degree = int(bool(dense))
    o.tier = params.tier
    o.eta = params.eta
    o.time_limit = params.time_limit
    return p, o


def _solve(instance: QuboProblem, params: Parameters, warm):
    L = _native.lib()
    n, m = instance.n(), instance.m()
    inst = _native.PraqInstance(n, m, _native.dptr(instance.diag), _native.iptr(instance.rows),
                                _native.iptr(instance.cols), _native.dptr(instance.vals))
    dense = instance.dense if params.dense_degree is None else params.dense_degree
    p, o = _native_params(L, params, dense)
    x = np.zeros(n)
    y = np.zeros((max(m, 1), 2))
    wx = wy = None
    if warm is not None:
        wx = np.ascontiguousarray(np.broadcast_to(np.asarray(warm[0], dtype=np.float64), (n,)))
        wy = _pairs_to_edges(instance, warm[1])
    info = _native.PraqInfo()
    status = L.praq_solve_ex(ctypes.byref(inst), ctypes.byref(p), ctypes.byref(o),
                             _native.dptr(wx), _native.dptr(wy), _native.dptr(x),
                             _native.dptr(y), ctypes.byref(info))
    if status < 0:
        raise ValueError(f"praq_solve_ex failed with status {status}")
    return x, y[:m], info


def solve_device(diag, rows, cols, vals, params: Parameters, x, y=None, stream=None):
    """One solve over device tensors (`praq_solve_device`): diag [n] and vals
    [m] float64, rows and cols [m] int64 (i < j), the answer written into the
    float64 tensors x [n] and y [m, 2] (y optional; dual in instance units),
    every launch and copy on `stream` (a torch stream; None: the current one).
    The sparse build (true degrees) unless params.dense_degree is set.
    Returns the PraqInfo record."""
    import torch
    L = _native.lib()
    n, m = int(diag.numel()), int(rows.numel())
    for name, t, dtype in (("diag", diag, torch.float64), ("rows", rows, torch.int64),
                           ("cols", cols, torch.int64), ("vals", vals, torch.float64),
                           ("x", x, torch.float64)):
        if not (t.is_cuda and t.dtype == dtype and t.is_contiguous()):
            raise ValueError(f"{name} must be a contiguous {dtype} device tensor")
    if y is not None and not (y.is_cuda and y.dtype == torch.float64 and y.is_contiguous()
                              and y.numel() == 2 * m):
        raise ValueError("y must be a contiguous float64 device tensor of shape [m, 2]")
    if x.numel() != n or cols.numel() != m or vals.numel() != m:
        raise ValueError("inconsistent sizes")
    p, o = _native_params(L, params, bool(params.dense_degree))
    inst = _native.PraqInstance(n, m, ctypes.cast(diag.data_ptr(), _native._dp),
                                ctypes.cast(rows.data_ptr(), _native._ip),
                                ctypes.cast(cols.data_ptr(), _native._ip),
                                ctypes.cast(vals.data_ptr(), _native._dp))
    s = stream if stream is not None else torch.cuda.current_stream(x.device)
    info = _native.PraqInfo()
    status = L.praq_solve_device(ctypes.byref(inst), ctypes.byref(p), ctypes.byref(o),
                                 x.data_ptr(), None if y is None else y.data_ptr(),
                                 ctypes.byref(info), s.cuda_stream)
    if status < 0:
        raise ValueError(f"praq_solve_device failed with status {status}: {_native.last_error()}")
    return info


def _pairs_to_edges(instance: QuboProblem, dual) -> np.ndarray:
    """A dual given per pair i < j (praq-jax's dense shape) or per edge."""
    d = np.asarray(dual, dtype=np.float64).reshape(-1, 2)
    n = instance.n()
    if instance.dense and d.shape[0] == n * (n - 1) // 2 and d.shape[0] != instance.m():
        # pair (i, j), i < j, at index i (2n - i - 1) / 2 + (j - i - 1)
        i, j = instance.rows, instance.cols
        idx = i * (2 * n - i - 1) // 2 + (j - i - 1)
        d = d[idx]
    return np.ascontiguousarray(d)


def _edges_to_pairs(instance: QuboProblem, y: np.ndarray) -> np.ndarray:
    if not instance.dense:
        return y
    n = instance.n()
    full = np.zeros((n * (n - 1) // 2, 2))
    i, j = instance.rows, instance.cols
    full[i * (2 * n - i - 1) // 2 + (j - i - 1)] = y
    return full


def solve_qubo_problem_warm_start(instance: QuboProblem, params: Parameters, warm_start):
    """(primal, dual, iterations) from a warm start (x, y) in instance units,
    projected before use (qubo/praq-jax/praq/solver.py:483-561)."""
    x, y, info = _solve(instance, params, warm_start)
    return x, _edges_to_pairs(instance, y), int(info.iterations)


def solve_qubo_problem(instance: QuboProblem, params: Parameters):
    """(primal, dual, iterations) from the centroid cold start
    (qubo/praq-jax/praq/solver.py:564-589)."""
    x, y, info = _solve(instance, params, None)
    return x, _edges_to_pairs(instance, y), int(info.iterations)


def solve_with_info(instance: QuboProblem, params: Parameters, warm_start=None):
    """The solve plus the solver's record (bounds, gap, restarts, tier)."""
    x, y, info = _solve(instance, params, warm_start)
    return x, _edges_to_pairs(instance, y), info


def relaxation_bounds(instance: QuboProblem, primal, dual):
    """(L_D, L_P) in instance units from any primal in [0,1]^n and dual in the
    dual set (qubo/praq-jax/praq/solver.py:348-362)."""
    x = np.asarray(primal, dtype=np.float64)
    y = _pairs_to_edges(instance, dual)
    i, j, w = instance.rows, instance.cols, 2.0 * np.abs(instance.vals)
    covering = instance.vals > 0.0
    xi, xj = x[i], x[j]
    edge = np.where(covering, w * np.maximum(xi + xj - 1.0, 0.0), -w * np.minimum(xi, xj))
    upper = float(instance.diag @ x + edge.sum())
    scat = np.zeros(instance.n())
    np.add.at(scat, i, y[:, 0])
    np.add.at(scat, j, y[:, 1])
    lower = float(np.minimum(instance.diag - scat, 0.0).sum() + y[covering, 0].sum())
    return lower, upper


def persistent_variables(primal, tolerance: float = 1e-6):
    """(mask of coordinates persistency fixes, their values)
    (qubo/praq-jax/praq/solver.py:592-611)."""
    x = np.asarray(primal, dtype=np.float64)
    rounded = np.round(x)
    return np.abs(x - rounded) <= tolerance, rounded


__all__ = ["Parameters", "QuboProblem", "check_interval", "solve_qubo_problem",
           "solve_qubo_problem_warm_start", "solve_with_info", "solve_device", "relaxation_bounds",
           "persistent_variables", "TIER_AUTO", "TIER_WORKGROUP", "TIER_DEVICE"]
