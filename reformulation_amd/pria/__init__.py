"""PRIA: the joint-replenishment LP relaxation, solved on MI355X.

The entry points, options and return values mirror pria-jax
(jrp/pria-jax/pria/prelude.py:1-15, solver.py:55-790, jrp_problem.py):
`JrpProblem`, `Parameters.with_auto_step`, `solve_jrp_problem`,
`solve_jrp_problem_warm_start`, `relaxation_bounds`, `primal_objective`.
Every solve runs the gfx950 kernels of libreformulation_hip.so through its
C ABI (`pria_solve_ex`).

Extensions: `Parameters` carries the iterate precision (`use_float32`), the
execution tier and a time limit.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from .. import _native

TIER_AUTO, TIER_WORKGROUP, TIER_DEVICE = 0, 1, 2


@dataclass(frozen=True)
class Parameters:
    """pria-jax's `Parameters` (jrp/pria-jax/pria/solver.py:55-252)."""

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
    iid: int = 0
    log: bool = False
    # extensions
    use_float32: bool = False
    tier: int = TIER_AUTO
    time_limit: float = 0.0

    @classmethod
    def new(cls, omega, eta, beta_sufficient, beta_artificial, beta_necessary, theta_initial,
            theta_target, theta_decay, max_iter, use_halpern, warm_steps, op_steps, tolerance,
            iid=0, log=False, **extensions) -> "Parameters":
        if max_iter is None:
            max_iter = -1
        return cls(float(omega), float(eta), float(beta_sufficient), float(beta_artificial),
                   float(beta_necessary), float(theta_initial), float(theta_target),
                   float(theta_decay), int(max_iter), bool(use_halpern), max(int(warm_steps), 0),
                   max(int(op_steps), 2), float(tolerance), iid, log, **extensions)

    @classmethod
    def with_auto_step(cls, max_iter, tolerance, use_halpern=True, op_steps=16, iid=0, log=False,
                       omega=0.10835, beta_sufficient=float(np.exp(-1.0)),
                       beta_necessary=0.5706, beta_artificial=0.2468, theta_initial=0.81930,
                       theta_target=0.42818, theta_decay=0.90474, warm_steps=0,
                       **extensions) -> "Parameters":
        eta = 0.999 / np.sqrt(2.0)
        return cls.new(omega, eta, beta_sufficient, beta_artificial, beta_necessary,
                       theta_initial, theta_target, theta_decay, max_iter, use_halpern,
                       warm_steps, op_steps, tolerance, iid, log, **extensions)

    @classmethod
    def stack(cls, parameters: list["Parameters"]) -> list["Parameters"]:
        return list(parameters)

    def native(self):
        p = _native.lib().pria_params_auto(int(self.max_iter), float(self.tolerance))
        p.omega, p.eta = self.omega, self.eta
        p.beta_sufficient, p.beta_artificial = self.beta_sufficient, self.beta_artificial
        p.beta_necessary = self.beta_necessary
        p.theta_initial, p.theta_target, p.theta_decay = (self.theta, self.theta_target,
                                                          self.theta_decay)
        p.use_halpern, p.warm_steps, p.op_steps = (int(self.use_halpern), int(self.warm_steps),
                                                   int(self.op_steps))
        return p


@dataclass(frozen=True)
class JrpProblem:
    """K_i = order_costs[i], H_ist = holding_costs[i][s][t] (upper triangle),
    K0 = joint_order_cost (jrp/pria-jax/pria/jrp_problem.py:12-68)."""

    order_costs: np.ndarray
    holding_costs: np.ndarray
    joint_order_cost: float

    def t(self) -> int:
        return int(self.holding_costs.shape[-1])

    def n(self) -> int:
        return int(self.order_costs.shape[0])

    @classmethod
    def new(cls, order_costs, joint_order_cost, holding_costs) -> "JrpProblem":
        h = np.triu(np.asarray(holding_costs, dtype=np.float64))
        return cls(np.asarray(order_costs, dtype=np.float64), h, float(joint_order_cost))

    @classmethod
    def stack(cls, instances: list["JrpProblem"]) -> list["JrpProblem"]:
        return list(instances)

    @classmethod
    def random_instance(cls, seed: int, items: int, time_periods: int) -> "JrpProblem":
        """Integer costs of the shape of pria-jax's generator
        (jrp_problem.py:84-113), drawn with numpy."""
        rng = np.random.default_rng(seed)
        t = time_periods
        diff = np.arange(t)[None, :] - np.arange(t)[:, None]
        mult = rng.uniform(0, 1, (items, 1, 1))
        add = rng.uniform(0, 1, (items, t, t)) * mult
        h = np.round(np.triu(mult * diff[None] + add, 1))
        allk = np.maximum(np.round(rng.uniform(1.0, max(t, 1.01), items + 1)), 1.0)
        return cls.new(allk[:-1], float(allk[-1]) if items > 1 else 0.0, h)

    @classmethod
    def zero_instance(cls, items: int, time_periods: int) -> "JrpProblem":
        return cls.new(np.zeros(items), 0.0, np.zeros((items, time_periods, time_periods)))

    def objective_value(self, primal) -> float:
        return float(primal_objective(self, primal))


def primal_objective(instance: JrpProblem, primal) -> float:
    """sum K_i max_t x + K0 sum_s max_{i,t} x + <H, x> (jrp_problem.py:143-161)."""
    x = np.asarray(primal, dtype=np.float64)
    return float(instance.order_costs @ x.max(axis=2).sum(axis=1)
                 + instance.joint_order_cost * x.max(axis=(0, 2)).sum()
                 + np.sum(x * instance.holding_costs))


def relaxation_bounds(instance: JrpProblem, primal, y, z):
    """(lower, upper) from any triple; the lower bound with the normalised
    duals scaled back by K_i, K0 (jrp/pria-jax/pria/solver.py:740-790)."""
    h = np.triu(instance.holding_costs)
    reduced = (h + instance.order_costs[:, None, None] * np.asarray(y)
               + instance.joint_order_cost * np.asarray(z))
    periods = np.arange(h.shape[-1])
    reduced = np.where(periods[:, None] <= periods[None, :], reduced, np.inf)
    return float(reduced.min(axis=1).sum()), primal_objective(instance, primal)


def _solve(instance: JrpProblem, params: Parameters, warm):
    L = _native.lib()
    n, t = instance.n(), instance.t()
    k = np.ascontiguousarray(instance.order_costs, dtype=np.float64)
    h = np.ascontiguousarray(instance.holding_costs, dtype=np.float64)
    p = params.native()
    o = L.pria_default_options()
    o.tier = params.tier
    o.time_limit = params.time_limit
    x, y, z = (np.zeros((n, t, t)) for _ in range(3))
    w = [None, None, None]
    if warm is not None:
        w = [np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (n, t, t)))
             for a in warm]
    info = _native.PriaInfo()
    status = L.pria_solve_ex(n, t, _native.dptr(k), float(instance.joint_order_cost),
                             _native.dptr(h), ctypes.byref(p), ctypes.byref(o),
                             int(params.use_float32), _native.dptr(w[0]), _native.dptr(w[1]),
                             _native.dptr(w[2]), _native.dptr(x), _native.dptr(y),
                             _native.dptr(z), ctypes.byref(info))
    if status < 0:
        raise ValueError(f"pria_solve_ex failed with status {status}")
    return x, y, z, info


def solve_device(order_costs, joint_order_cost: float, holding_costs, params: Parameters,
                 x=None, y=None, z=None, stream=None):
    """One cold-start solve over device tensors (`pria_solve_device`):
    order_costs [n] and holding_costs [n, t, t] float64 (upper triangle
    read); the answer written into the float64 tensors x, y, z [n, t, t] (each
    optional; y, z normalised as solve_jrp_problem returns them), every launch
    and copy on `stream` (a torch stream; None: the current one).  Returns the
    PriaInfo record."""
    import torch
    L = _native.lib()
    if holding_costs.dim() != 3 or holding_costs.shape[1] != holding_costs.shape[2]:
        raise ValueError("holding_costs must have shape [n, t, t]")
    n, t = int(holding_costs.shape[0]), int(holding_costs.shape[1])
    for name, a, count in (("order_costs", order_costs, n), ("holding_costs", holding_costs,
                                                               n * t * t),
                           ("x", x, n * t * t), ("y", y, n * t * t), ("z", z, n * t * t)):
        if a is None:
            continue
        if not (a.is_cuda and a.dtype == torch.float64 and a.is_contiguous()
                and a.numel() == count):
            raise ValueError(f"{name} must be a contiguous float64 device tensor of {count}")
    p = params.native()
    o = L.pria_default_options()
    o.tier = params.tier
    o.time_limit = params.time_limit
    s = stream if stream is not None else torch.cuda.current_stream(holding_costs.device)
    ptr = [None if a is None else a.data_ptr() for a in (x, y, z)]
    info = _native.PriaInfo()
    status = L.pria_solve_device(n, t, order_costs.data_ptr(), float(joint_order_cost),
                                 holding_costs.data_ptr(), ctypes.byref(p), ctypes.byref(o),
                                 int(params.use_float32), ptr[0], ptr[1], ptr[2],
                                 ctypes.byref(info), s.cuda_stream)
    if status < 0:
        raise ValueError(f"pria_solve_device failed with status {status}: {_native.last_error()}")
    return info


def solve_jrp_problem_warm_start(instance: JrpProblem, params: Parameters, warm_start):
    """(x, y, z, iterations, lower, upper) from a warm start whose duals are
    normalised as the solver returns them (solver.py:608-724)."""
    x, y, z, info = _solve(instance, params, warm_start)
    return x, y, z, int(info.iterations), float(info.lower), float(info.upper)


def solve_jrp_problem(instance: JrpProblem, params: Parameters):
    """(x, y, z, iterations, lower, upper) from zeros (solver.py:727-737)."""
    x, y, z, info = _solve(instance, params, None)
    return x, y, z, int(info.iterations), float(info.lower), float(info.upper)


def solve_with_info(instance: JrpProblem, params: Parameters, warm_start=None):
    return _solve(instance, params, warm_start)


__all__ = ["Parameters", "JrpProblem", "solve_jrp_problem", "solve_jrp_problem_warm_start",
           "solve_with_info", "solve_device", "relaxation_bounds", "primal_objective", "TIER_AUTO",
           "TIER_WORKGROUP", "TIER_DEVICE"]
