"""praq-hip: PRAQ on MI355X as a harness backend.

The contract of qubo/harness/backends/praq_cuda.py:79-167 (status mapping,
objective = upper only when optimal, bracket always reported) without its
subprocess: the solve runs in-process through the C ABI, as the reference's
tuner drives libpraq.so (qubo/harness/praq_cuda_library.py:123-139).
`time_limit` is enforced by the solver itself (its wall clock stops the
iteration and the bracket reached is reported as `time_limit`).
"""
from __future__ import annotations

import math
import time

from ... import praq
from ..results import memory_fields, result_class

#: statuses of the solver record -> harness statuses
_STATUS = {0: "optimal", 1: "iteration_limit", 2: "time_limit"}


def solve(instance, *, tolerances, time_limit: float, use_float32: bool = False,
          sample_memory: bool = False):
    Result = result_class("qubo")
    common = dict(backend="praq-hip", instance=instance.name, family=instance.family,
                  n=int(instance.n), edges=int(len(instance.vals)),
                  optimality_tolerance=float(tolerances.optimality))
    try:
        start = time.perf_counter()
        problem = praq.QuboProblem.from_instance(instance)
        params = praq.Parameters.with_auto_step(None, tolerances.optimality,
                                                use_float32=use_float32,
                                                time_limit=float(time_limit))
        setup = time.perf_counter() - start
        start = time.perf_counter()
        x, y, info = praq.solve_with_info(problem, params)
        elapsed = time.perf_counter() - start
    except Exception as exc:  # reported, as the harness expects, not raised
        return Result(status="error", note=str(exc)[:200], **common)
    status = _STATUS.get(int(info.status), "error")
    return Result(status=status, objective=float(info.upper) if status == "optimal" else math.nan,
                  lower=float(info.lower), upper=float(info.upper),
                  iterations=int(info.iterations), setup_seconds=setup, solve_seconds=elapsed,
                  **memory_fields(info, sample_memory), **common)
