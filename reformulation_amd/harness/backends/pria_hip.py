"""pria-hip: PRIA on MI355X as a harness backend.

The contract of jrp/harness/backends/pria_cuda.py:83-173, in-process through
the C ABI; `time_limit` is enforced by the solver.
"""
from __future__ import annotations

import math
import time

from ... import pria
from ..results import memory_fields, result_class

_STATUS = {0: "optimal", 1: "iteration_limit", 2: "time_limit"}


def solve(instance, *, tolerances, time_limit: float, use_float32: bool = False,
          sample_memory: bool = False):
    Result = result_class("jrp")
    items, periods = int(instance.order_costs.size), int(instance.holding_costs.shape[-1])
    common = dict(backend="pria-hip", instance=instance.name, family=instance.family,
                  items=items, periods=periods, entries=items * periods * (periods + 1) // 2,
                  optimality_tolerance=float(tolerances.optimality))
    try:
        start = time.perf_counter()
        problem = pria.JrpProblem.new(instance.order_costs, instance.joint_order_cost,
                                      instance.holding_costs)
        params = pria.Parameters.with_auto_step(None, tolerances.optimality,
                                                use_float32=use_float32,
                                                time_limit=float(time_limit))
        setup = time.perf_counter() - start
        start = time.perf_counter()
        x, y, z, info = pria.solve_with_info(problem, params)
        elapsed = time.perf_counter() - start
    except Exception as exc:
        return Result(status="error", note=str(exc)[:200], **common)
    status = _STATUS.get(int(info.status), "error")
    return Result(status=status, objective=float(info.upper) if status == "optimal" else math.nan,
                  lower=float(info.lower), upper=float(info.upper),
                  iterations=int(info.iterations), setup_seconds=setup, solve_seconds=elapsed,
                  **memory_fields(info, sample_memory), **common)
