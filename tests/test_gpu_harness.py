"""The harness backends (reformulation_amd.harness.REGISTRY) with the
reference harness's calling convention: solve(instance, *, tolerances,
time_limit, sample_memory) -> Result (qubo/harness/results.py:24-80,
jrp/harness/results.py:25-70).  Marked gpu.

Bar: status "optimal" with objective = upper, the bracket holding the LP
value (HiGHS), and, with sample_memory, the device bytes the library held."""
import importlib
from types import SimpleNamespace

import numpy as np
import pytest

from lp_models import praq_lp_value, pria_lp_value

pytestmark = pytest.mark.gpu


def _backend(name):
    from reformulation_amd.harness import REGISTRY
    return importlib.import_module(REGISTRY[name])


def test_praq_backend_reports_the_reference_record():
    from reformulation_amd import workloads
    w = workloads.praq_er(3, 2000, 8000)
    inst = SimpleNamespace(name="er-2000-3", family="erdos-renyi", n=w.n, diag=w.diag,
                           rows=w.rows, cols=w.cols, vals=w.vals)
    tol = SimpleNamespace(optimality=1e-6, feasibility=float("nan"))
    r = _backend("praq-hip").solve(inst, tolerances=tol, time_limit=60.0, sample_memory=True)
    assert r.status == "optimal" and r.objective == r.upper
    v = praq_lp_value(inst)
    slack = 1e-9 * (1 + abs(v))
    assert r.lower <= v + slack and r.upper >= v - slack
    assert (r.upper - r.lower) <= 1e-6 * (1 + abs(r.upper) + abs(r.lower))
    assert r.iterations > 0 and r.solve_seconds > 0 and r.peak_device_bytes > 0
    assert r.peak_host_bytes > 0


def test_pria_backend_reports_the_reference_record_and_time_limit():
    from reformulation_amd import workloads
    j = workloads.pria_c1(5, 12, 41)
    inst = SimpleNamespace(name="c1-12-41", family="c1", order_costs=np.asarray(j.order_costs),
                           joint_order_cost=float(j.joint_order_cost),
                           holding_costs=np.asarray(j.holding_costs))
    tol = SimpleNamespace(optimality=1e-6, feasibility=float("nan"))
    r = _backend("pria-hip").solve(inst, tolerances=tol, time_limit=60.0, sample_memory=True)
    assert r.status == "optimal" and r.objective == r.upper
    v = pria_lp_value(inst)
    slack = 1e-9 * (1 + abs(v))
    assert r.lower <= v + slack and r.upper >= v - slack
    assert r.peak_device_bytes > 0
    # a limit the solve cannot meet: the bracket so far, status time_limit
    big = workloads.pria_c1(6, 50, 200)
    inst2 = SimpleNamespace(name="c1-50-200", family="c1",
                            order_costs=np.asarray(big.order_costs),
                            joint_order_cost=float(big.joint_order_cost),
                            holding_costs=np.asarray(big.holding_costs))
    tight = SimpleNamespace(optimality=1e-13, feasibility=float("nan"))
    r2 = _backend("pria-hip").solve(inst2, tolerances=tight, time_limit=0.05)
    assert r2.status == "time_limit" and np.isnan(r2.objective)
    assert np.isfinite(r2.lower) and np.isfinite(r2.upper) and r2.lower <= r2.upper + 1e-6
