"""PRIA on the GPU against the oracle and the reference CPU solver's outputs,
through the C ABI.  Marked gpu.

Parity bar (FP64 iterates): iterations equal to the oracle's (which equal the
reference CPU solver's, tests/golden/pria_ref.npz); bounds within 1e-9
relative; x, y, z within 1e-8 (simplex thresholds found by Michelot on the
GPU and by sorting on the CPU agree to rounding).  FP32 iterates: the bracket
closes to the requested gap around the LP value."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from helpers import small_jrp

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def _gpu(k, k0, h, tol, tier=0, f32=False, max_iter=None, warm=None):
    from reformulation_amd import pria
    problem = pria.JrpProblem.new(k, k0, h)
    params = pria.Parameters.with_auto_step(max_iter, tol, use_float32=f32, tier=tier,
                                            time_limit=120.0)
    return pria.solve_with_info(problem, params, warm)


def _ref_cases():
    data = np.load(GOLDEN / "pria_ref.npz")
    for q in range(len(data["cases"])):
        yield q, {k.split("_", 1)[1]: data[k] for k in data.files if k.startswith(f"c{q}_")}


@pytest.mark.parametrize("tier", [1, 2])
def test_fp64_matches_the_reference_cpu_solver(tier):
    for q, c in _ref_cases():
        x, y, z, info = _gpu(c["k"], float(c["k0"]), c["h"], 1e-8, tier=tier)
        assert info.tier == tier
        assert info.iterations == int(c["iterations"]), q
        sc = 1.0 + abs(float(c["primal"]))
        assert abs(info.upper_iterate - float(c["primal"])) <= 1e-9 * sc, q
        assert abs(info.lower - float(c["dual"])) <= 1e-9 * sc, q
        assert np.max(np.abs(x - c["x"])) <= 1e-8, q
        assert np.max(np.abs(y - c["y"])) <= 1e-8, q
        assert np.max(np.abs(z - c["z"])) <= 1e-8, q


@pytest.mark.parametrize("tier", [1, 2])
def test_fp64_matches_the_oracle(tier):
    for seed, n, t in [(1, 3, 6), (2, 4, 7), (8, 12, 33), (9, 1, 5), (10, 20, 24)]:
        k, k0, h = small_jrp(seed, n, t)
        o = oracle.pria_solve(k, k0, h, oracle.pria_params(tolerance=1e-8))
        x, y, z, info = _gpu(k, k0, h, 1e-8, tier=tier)
        assert info.iterations == o.iterations and info.restarts == o.restarts, (seed, n, t)
        sc = 1.0 + abs(o.upper)
        assert abs(info.upper - o.upper) <= 1e-9 * sc
        assert abs(info.lower - o.lower) <= 1e-9 * sc
        assert np.max(np.abs(x - o.x)) <= 1e-8
        assert np.max(np.abs(y - o.y)) <= 1e-8
        assert np.max(np.abs(z - o.z)) <= 1e-8


def test_highs_values_and_certificates():
    from reformulation_amd import pria, workloads
    rows = json.loads((GOLDEN / "pria_lp.json").read_text())
    for row in rows:
        inst = workloads.pria_c1(row["seed"], row["items"], row["periods"])
        for f32 in (False, True):
            x, y, z, info = _gpu(inst.order_costs, inst.joint_order_cost, inst.holding_costs,
                                 1e-6 if f32 else 1e-9, f32=f32)
            opt = row["highs"]
            tol = 2e-6 if f32 else 1e-7
            assert abs(info.upper - opt) <= tol * (1 + abs(opt))
            lo, up = pria.relaxation_bounds(inst.problem(), x, y, z)
            assert lo <= opt + tol * abs(opt) and up >= opt - tol * abs(opt)


def test_special_cases():
    # K0 = 0 gives z = 0; max_iter = 64 gives exactly 64; a returned solution
    # warm-starts as a fixed point at even T
    k, k0, h = small_jrp(3, 1, 8)
    x, y, z, info = _gpu(k, 0.0, h, 1e-8)
    assert np.all(z == 0.0)
    k, k0, h = small_jrp(113, 6, 18)
    x, y, z, info = _gpu(k, k0, h, 1e-14, max_iter=64)
    assert info.iterations == 64 and info.status == 1
    k, k0, h = small_jrp(4, 3, 8)
    x, y, z, info = _gpu(k, k0, h, 1e-7)
    x2, y2, z2, info2 = _gpu(k, k0, h, 1e-7, warm=(x, y, z))
    assert info2.iterations == 0


def test_plain_config_fp64_against_oracle():
    """The plain-PRIA workload (N = 50, T = 200) on the device tier, FP64."""
    from reformulation_amd import workloads
    inst = workloads.pria_c1(0, 50, 200)
    x, y, z, info = _gpu(inst.order_costs, inst.joint_order_cost, inst.holding_costs, 1e-6,
                         tier=2, max_iter=50_000)
    assert info.status == 0
    gold = json.loads((GOLDEN / "pria_c1.json").read_text())
    assert info.iterations == gold["iterations"]
    assert abs(info.upper - gold["upper"]) <= 1e-9 * (1 + abs(gold["upper"]))
    assert abs(info.lower - gold["lower"]) <= 1e-9 * (1 + abs(gold["lower"]))
    # the bracket holds HiGHS's value of the full explicit LP
    v, slack = gold["highs"], 1e-8 * (1 + abs(gold["highs"]))
    assert info.lower <= v + slack and info.upper >= v - slack
    # FP32 iterates (FP64 finish) too
    _, _, _, i32 = _gpu(inst.order_costs, inst.joint_order_cost, inst.holding_costs, 1e-6,
                        tier=2, f32=True, max_iter=50_000)
    assert i32.status == 0 and i32.lower <= v + slack and i32.upper >= v - slack


def test_concurrent_solves_from_threads_equal_sequential_ones():
    """pria_solve_ex on several host threads at once (each solve on its own
    stream, stream-ordered allocations): every answer is bitwise the one the
    same solve gives alone."""
    import threading
    import time
    from reformulation_amd import pria, workloads
    cases = [workloads.pria_c1(300 + q, n, t) for q, (n, t) in
             enumerate([(20, 60), (30, 90), (8, 17), (40, 120), (12, 41), (25, 80)])]
    params = pria.Parameters.with_auto_step(100_000, 1e-6)
    alone = []
    t0 = time.perf_counter()
    for c in cases:
        alone.append(pria.solve_with_info(c.problem(), params))
    t_seq = time.perf_counter() - t0
    together = [None] * len(cases)

    def work(q):
        together[q] = pria.solve_with_info(cases[q].problem(), params)

    threads = [threading.Thread(target=work, args=(q,)) for q in range(len(cases))]
    t0 = time.perf_counter()
    for th in threads:
        th.start()
    for th in threads:
        th.join(120.0)
    t_par = time.perf_counter() - t0
    print(f"pria threads: sequential {t_seq:.3f} s, {len(cases)} threads {t_par:.3f} s", flush=True)
    for q, (a, b) in enumerate(zip(alone, together)):
        assert b is not None, q
        for u, v in zip(a[:3], b[:3]):
            assert np.array_equal(u, v), q
        assert a[3].iterations == b[3].iterations and a[3].upper == b[3].upper, q


@pytest.mark.parametrize("tier", [1, 2])
def test_designed_instances_follow_the_oracle_and_hit_their_values(tier):
    """The reference's designed instances (fractional optima of denominators
    3..34, the p = 8 tie and its hardest bump set; jrp/pria-jax/pria/
    adversarial.py): FP64 iterations and restarts equal the oracle's, bounds
    agree to 1e-9, and the bracket holds the designed LP value (item 1's
    plus 2 T); FP32 iterates (FP64 finish) reach the same value to 1e-6."""
    import adversarial as A
    cases = [(f"fractional p={p}", A.fractional_optimum(p)) for p in (4, 6, 7, 8)]
    cases += [("tie p=8", A.tie_split(())), ("tie p=8 hard bumps", A.tie_split())]
    for name, (k, k0, h, info) in cases:
        value = A.single_item_lp(info["item"], info["K"])[0] + 2 * info["T"]
        print(f"designed {name} tier {tier}", flush=True)
        o = oracle.pria_solve(k, k0, h, oracle.pria_params(tolerance=1e-9, max_iter=200_000))
        x, y, z, g = _gpu(k, k0, h, 1e-9, tier=tier, max_iter=200_000)
        assert g.status == 0 and g.tier == tier, name
        assert g.iterations == o.iterations and g.restarts == o.restarts, name
        sc = 1.0 + abs(o.upper)
        assert abs(g.upper - o.upper) <= 1e-9 * sc and abs(g.lower - o.lower) <= 1e-9 * sc, name
        tol = 1e-9 * (1 + abs(g.upper) + abs(g.lower))
        assert g.lower <= value + tol and g.upper >= value - tol, name
        _, _, _, g32 = _gpu(k, k0, h, 1e-6, tier=tier, f32=True, max_iter=200_000)
        assert g32.status == 0 and abs(g32.upper - value) <= 2e-6 * (1 + abs(value)), name


def test_designed_instance_with_the_reference_smoke_schedule():
    """jrp/pria-jax/tests/test_adversarial.py:80-97 on the GPU: beta_necessary
    0, beta_artificial 10, theta 0 at 1e-9; iterations as the oracle's."""
    import adversarial as A
    from reformulation_amd import pria
    k, k0, h, _ = A.fractional_optimum(4)
    o = oracle.pria_solve(k, k0, h, oracle.pria_params(
        tolerance=1e-9, max_iter=50_000, beta_necessary=0.0, beta_artificial=10.0,
        theta_initial=0.0, theta_target=0.0))
    params = pria.Parameters.with_auto_step(50_000, 1e-9, beta_necessary=0.0,
                                            beta_artificial=10.0, theta_initial=0.0,
                                            theta_target=0.0)
    x, y, z, info = pria.solve_with_info(pria.JrpProblem.new(k, k0, h), params)
    assert 0 < info.iterations < 50_000 and info.iterations == o.iterations
    assert all(np.isfinite(a).all() for a in (x, y, z))
