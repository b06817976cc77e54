"""The precision statement every result carries: switch_iteration and the
bracket U - L (include/praq_hip.h PraqInfo, include/pria_hip.h pria_info).
Marked gpu.

Against LP optima fixed independently of this solver (HiGHS on the explicit
LP, and the max-flow value of the roof dual, tests/golden/praq_lp.json,
pria_lp.json, pria_c1.json):

* bracket is exactly upper - lower, and a converged solve's relative gap
  (U - L) / (1 + |U| + |L|) is within the tolerance;
* a solve that finished in FP64 (FP64 iterates, or FP32 ones that handed
  over, switch_iteration >= 0) brackets the optimum: L <= OPT <= U to FP64
  rounding, so the answer's objective error U - OPT is at most the bracket;
* one that finished in FP32 (switch_iteration = -1) brackets the optimum of
  the instance its FP32 image holds, costs rounded to FP32: L and U within
  2^-24 times the cost mass of OPT (PRIA's U is evaluated on the original
  instance at an x feasible to FP32 rounding);
* FP32 iterates with a tolerance below 3e-5 always finish in FP64;
* switch_iteration is -1 for FP64 iterates and lies in [0, iterations] when
  FP32 iterates handed over."""
import json
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).parent / "golden"
U24 = 2.0 ** -24


def _check(info, opt, tol, fp32, eps):
    """The statement for one converged result; True if it finished in FP64.
    PRIA's termination test used upper_iterate (the padded instance as the
    image holds it), at least the original objective U up to rounding."""
    assert info.status == 0 and info.relative_gap <= tol
    assert info.bracket == info.upper - info.lower
    slack = 1e-12 * (1.0 + abs(opt))          # FP64 rounding of the sums
    sw = info.switch_iteration
    if not fp32:
        assert sw == -1
    else:
        assert sw == -1 or 0 <= sw <= info.iterations
    in64 = not fp32 or sw >= 0
    if tol < 3e-5:
        assert in64, "FP32 iterates below the switch gap finish in FP64"
    room = slack if in64 else eps + slack
    u_test = getattr(info, "upper_iterate", info.upper)
    assert info.bracket <= tol * (1 + abs(u_test) + abs(info.lower)) * (1 + 1e-9) + room
    assert info.lower <= opt + room
    assert info.upper >= opt - room
    return in64


def test_praq_brackets_hold_the_optimum():
    from reformulation_amd import praq, workloads
    rows = json.loads((GOLDEN / "praq_lp.json").read_text())
    finishes = set()
    for row in rows[:2]:
        inst = workloads.praq_er(row["seed"], row["n"], row["m"])
        problem = inst.problem()
        opt = row["highs"]
        assert abs(row["maxflow"] - opt) <= 1e-9 * abs(opt)
        eps = U24 * (np.abs(inst.diag).sum() + 2.0 * np.abs(inst.vals).sum())
        for fp32, tol in ((False, 1e-6), (True, 1e-6), (True, 1e-3)):
            print(f"praq n={row['n']} fp32={fp32} tol={tol:g}", flush=True)
            params = praq.Parameters.with_auto_step(100_000, tol, use_float32=fp32)
            _, _, info = praq.solve_with_info(problem, params)
            finishes.add((fp32, _check(info, opt, tol, fp32, eps)))
    # both kinds of FP32 finish were exercised
    assert (True, True) in finishes and (True, False) in finishes


@pytest.mark.parametrize("which", ["lp", "c1"])
def test_pria_brackets_hold_the_optimum(which):
    from reformulation_amd import pria, workloads
    if which == "lp":
        cases = [(r["seed"], r["items"], r["periods"], r["highs"])
                 for r in json.loads((GOLDEN / "pria_lp.json").read_text())]
    else:
        gold = json.loads((GOLDEN / "pria_c1.json").read_text())
        cases = [(0, 50, 200, gold["highs"])]
    finishes = set()
    for seed, n, t, opt in cases:
        inst = workloads.pria_c1(seed, n, t)
        problem = inst.problem()
        h = inst.holding_costs
        mass = sum(np.abs(np.triu(h[i])).sum() for i in range(n))
        eps = U24 * (mass + t * np.abs(inst.order_costs).sum() + t * abs(inst.joint_order_cost))
        for fp32, tol in ((False, 1e-6), (True, 1e-6), (True, 1e-3)):
            print(f"pria {n}x{t} fp32={fp32} tol={tol:g}", flush=True)
            params = pria.Parameters.with_auto_step(100_000, tol, use_float32=fp32)
            _, _, _, info = pria.solve_with_info(problem, params)
            finishes.add((fp32, _check(info, opt, tol, fp32, eps)))
    assert (True, False) in finishes


def test_batched_and_served_results_carry_the_bracket():
    from reformulation_amd import praq, pria, workloads
    from reformulation_amd.batch import Batch
    from reformulation_amd.serve import Server
    q = workloads.praq_er(0, 2000, 8000).problem()
    j = workloads.pria_c1(3, 12, 41).problem()
    pp = praq.Parameters.with_auto_step(100_000, 1e-6)
    rp = pria.Parameters.with_auto_step(100_000, 1e-6)
    b = Batch(use_float32=True)
    iq, ij = b.add_praq(q, pp), b.add_pria(j, rp)
    b.upload()
    b.solve("batched")
    infos = [b.praq_result(iq)[2], b.pria_result(ij)[3]]
    b.close()
    with Server(use_float32=True, host_threads=2) as srv:
        tickets = [srv.submit_praq(q, pp), srv.submit_pria(j, rp)]
        srv.wait_idle(60.0)
        infos += [srv.info(t) for t in tickets]
    for info in infos:
        assert info.status == 0
        assert info.bracket == info.upper - info.lower
        # within the tolerance, give or take FP32 rounding of the data
        assert info.bracket <= 1.1e-6 * (1 + abs(info.upper) + abs(info.lower))
