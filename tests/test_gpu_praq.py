"""PRAQ on the GPU against the oracle, through the C ABI.  Marked gpu.

Parity bar (FP64 iterates): the same iteration and restart counts as the
oracle, bounds within 1e-9 relative, primal within 1e-8 absolute -- the
arithmetic is the same up to summation order.  FP32 iterates: the bracket
holds the LP value and closes to the requested gap."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from helpers import check_answer_properties, half_integral_optimum, random_qubo, relaxation_value

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def _gpu(inst, tol, tier=0, f32=False, max_iter=None, dense=False, warm=None):
    from reformulation_amd import praq
    n, d, r, c, v = inst
    problem = praq.QuboProblem.from_sparse(n, d, r, c, v)
    params = praq.Parameters.with_auto_step(max_iter, tol, use_float32=f32, tier=tier,
                                            dense_degree=dense, time_limit=60.0)
    return praq.solve_with_info(problem, params, warm)


@pytest.mark.parametrize("tier", [1, 2])
def test_fp64_matches_the_oracle_step_for_step(tier):
    seed = 0
    for n in range(1, 9):
        for density in (1.0, 0.5, 0.2, 0.0):
            inst = random_qubo(seed, n, density)
            seed += 1
            for dense in (False, True):
                o = oracle.praq_solve(*inst, params=oracle.praq_params(tolerance=1e-9,
                                                                       dense_degree=dense))
                x, y, info = _gpu(inst, 1e-9, tier=tier, dense=dense)
                assert info.tier == tier
                assert info.iterations == o.iterations and info.restarts == o.restarts, (n, density)
                sc = 1.0 + abs(o.upper)
                assert abs(info.upper - o.upper) <= 1e-9 * sc
                assert abs(info.lower - o.lower) <= 1e-9 * sc
                assert np.max(np.abs(x - o.x)) <= 1e-8
                if len(inst[4]):
                    assert np.max(np.abs(y - o.y)) <= 1e-8 * sc


def test_matches_enumeration_and_certifies_itself():
    seed = 500
    for n in range(2, 8):
        inst = random_qubo(seed, n, 0.6)
        seed += 1
        best = half_integral_optimum(*inst)
        x, y, info = _gpu(inst, 1e-9)
        sc = 1.0 + abs(best)
        assert info.status == 0
        assert info.lower <= best + 1e-9 * sc and info.upper >= best - 1e-9 * sc
        assert abs(relaxation_value(*inst, x) - info.upper) <= 1e-9 * sc


@pytest.mark.parametrize("tier", [1, 2])
def test_config_size_fp64_matches_oracle_highs_and_maxflow(tier):
    from reformulation_amd import workloads
    rows = json.loads((GOLDEN / "praq_lp.json").read_text())
    for row in rows[:2]:
        inst = workloads.praq_er(row["seed"], row["n"], row["m"])
        tup = (inst.n, inst.diag, inst.rows, inst.cols, inst.vals)
        o = oracle.praq_solve(*tup, params=oracle.praq_params(tolerance=1e-6))
        x, y, info = _gpu(tup, 1e-6, tier=tier)
        assert info.status == 0
        assert info.iterations == o.iterations
        opt = row["highs"]
        assert abs(row["maxflow"] - opt) <= 1e-9 * abs(opt)
        assert info.lower <= opt + 1e-9 * abs(opt) <= info.upper + 2e-9 * abs(opt)
        assert info.upper - opt <= 1e-6 * (1 + abs(info.upper) + abs(info.lower))
        assert abs(info.upper - o.upper) <= 1e-9 * (1 + abs(o.upper))


@pytest.mark.parametrize("tier", [1, 2])
def test_fp32_iterates_reach_1e6(tier):
    from reformulation_amd import workloads
    rows = json.loads((GOLDEN / "praq_lp.json").read_text())
    for row in rows[:2]:
        inst = workloads.praq_er(row["seed"], row["n"], row["m"])
        tup = (inst.n, inst.diag, inst.rows, inst.cols, inst.vals)
        x, y, info = _gpu(tup, 1e-6, tier=tier, f32=True, max_iter=100_000)
        assert info.status == 0
        opt = row["highs"]
        # the bracket is computed in FP64 from FP32 iterates: allow their rounding
        assert info.lower <= opt + 1e-6 * abs(opt) and opt <= info.upper + 1e-6 * abs(opt)
        assert abs(info.upper - opt) <= 2e-6 * (1 + abs(opt))


def test_bounds_bracket_at_every_budget():
    inst = random_qubo(101, 8, 0.6)
    best = half_integral_optimum(*inst)
    sc = 1.0 + abs(best)
    for budget in (1, 4, 16, 64, 256):
        x, y, info = _gpu(inst, 1e-12, max_iter=budget)
        assert info.iterations <= budget + 15
        assert info.lower <= best + 1e-9 * sc and info.upper >= best - 1e-9 * sc
        assert info.status == 1 or info.relative_gap <= 1e-12


def test_warm_start_and_special_cases():
    inst = random_qubo(5, 7, 0.6)
    x, y, info = _gpu(inst, 1e-9)
    x2, y2, info2 = _gpu(inst, 1e-9, warm=(x, y))
    assert info2.iterations <= 16
    # n = 1
    _, _, i1 = _gpu((1, np.array([-3.0]), np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0)), 1e-9)
    assert abs(i1.upper + 3.0) < 1e-8
    # max-cut form ends at iteration 0
    n, _, rows, cols, vals = random_qubo(7, 9, 0.7)
    vals = np.abs(vals)
    diag = np.zeros(n)
    np.add.at(diag, rows, -vals)
    np.add.at(diag, cols, -vals)
    _, _, im = _gpu((n, diag, rows, cols, vals), 1e-9)
    assert im.iterations == 0


def test_vanilla_and_warm_phase_options_follow_the_oracle():
    inst = random_qubo(42, 8, 0.7)
    from reformulation_amd import praq
    problem = praq.QuboProblem.from_sparse(*inst)
    for kw in (dict(use_halpern=False, max_iter=400), dict(warm_steps=2)):
        params = praq.Parameters.with_auto_step(kw.get("max_iter"), 1e-9,
                                                use_halpern=kw.get("use_halpern", True),
                                                warm_steps=kw.get("warm_steps", 0))
        x, y, info = praq.solve_with_info(problem, params)
        o = oracle.praq_solve(*inst, params=oracle.praq_params(
            tolerance=1e-9, max_iter=kw.get("max_iter", -1), use_halpern=kw.get("use_halpern", True),
            warm_steps=kw.get("warm_steps", 0)))
        assert info.iterations == o.iterations
        assert abs(info.upper - o.upper) <= 1e-9 * (1 + abs(o.upper))


def _hub_instance(seed=41, n=3000):
    """Two hubs (degrees n - 1 and n / 2, several chunks each) over an
    Erdos-Renyi graph of mean degree 4, integer weights."""
    from reformulation_amd import workloads
    base = workloads.praq_er(seed, n, 2 * n)
    rng = np.random.default_rng(seed)
    pairs = {(int(i), int(j)) for i, j in zip(base.rows, base.cols)}
    pairs |= {(0, j) for j in range(1, n)}
    pairs |= {(1, int(j)) for j in rng.choice(np.arange(2, n), n // 2, replace=False)}
    rows, cols = (np.array(v, dtype=np.int64) for v in zip(*sorted(pairs)))
    vals = 0.5 * (rng.integers(1, 101, rows.size) * rng.choice([-1, 1], rows.size))
    return n, base.diag, rows, cols, vals.astype(float)


def test_hub_nodes_by_chunks_match_the_oracle_and_the_workgroup_tier():
    """Nodes of degree above 64 lanes take the chunked gather in the device
    tier (one wavefront per 1024 half-edges, FP64 chunk sums in order): FP64
    results keep the oracle's iteration counts and bounds to 1e-9, and the
    workgroup tier (which walks hubs with its lane groups) agrees."""
    inst = _hub_instance()
    o = oracle.praq_solve(*inst, params=oracle.praq_params(tolerance=1e-7))
    sc = 1.0 + abs(o.upper)
    print(f"hub instance: n={inst[0]} m={len(inst[4])}, oracle {o.iterations} iterations",
          flush=True)
    for tier in (2, 1):
        x, y, info = _gpu(inst, 1e-7, tier=tier)
        assert info.tier == tier and info.status == 0
        assert info.iterations == o.iterations and info.restarts == o.restarts, tier
        assert abs(info.upper - o.upper) <= 1e-9 * sc and abs(info.lower - o.lower) <= 1e-9 * sc
        assert np.max(np.abs(x - o.x)) <= 1e-7
    x, y, info = _gpu(inst, 1e-6, tier=2, f32=True, max_iter=200_000)
    assert info.status == 0 and abs(info.upper - o.upper) <= 2e-6 * sc



@pytest.mark.parametrize("tier", [1, 2])
def test_persistency_and_dual_feasibility_of_the_answer(tier):
    """qubo/praq-jax/tests/test_solver.py:71-134 on the GPU's answers (the
    checks in helpers.check_answer_properties)."""
    for seed, n in [(11, 4), (12, 6), (13, 8), (14, 10), (7, 9)]:
        inst = random_qubo(seed, n, 0.6)
        x, y, info = _gpu(inst, 1e-9, tier=tier, max_iter=200_000)
        assert info.status == 0, seed
        check_answer_properties(inst, x, y)


def test_large_chung_lu_bracket_holds_the_max_flow_value():
    """The large instance (n = 1e6, m = 1.6e7, Chung-Lu degrees, Gaussian
    weights; device tier with hub chunks), FP64 and FP32 iterates at 1e-6:
    the final bracket L <= OPT <= U holds around the relaxation value that
    the oracle's int64 max-flow roof dual gives (tests/golden/
    c2_roof_dual.json, to its capacity-rounding bound), so both ends are
    within the requested gap of the optimum."""
    from reformulation_amd import praq, workloads
    ref = json.loads((GOLDEN / "c2_roof_dual.json").read_text())
    opt, slack = ref["roof_dual"], ref["error_bound"]
    w = workloads.praq_chung_lu(0)
    assert (w.n, len(w.vals)) == (ref["n"], ref["m"])
    problem = w.problem()
    for f32 in (False, True):
        print(f"chung-lu 1e6 f32={f32}", flush=True)
        params = praq.Parameters.with_auto_step(100_000, 1e-6, use_float32=f32, time_limit=120.0)
        _, _, info = praq.solve_with_info(problem, params)
        assert info.status == 0, f32
        assert info.lower <= opt + slack and info.upper >= opt - slack, (f32, info.lower, info.upper)
        gap = 1e-6 * (1 + abs(info.upper) + abs(info.lower))
        assert info.upper - opt <= gap + slack and opt - info.lower <= gap + slack, f32
