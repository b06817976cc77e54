"""Parity at the sizes the measured workloads run, through the C ABI.  Marked gpu.

  * PRIA against the reference's own CPU solver (jrp/pria-cuda/tests/ref_impl.hpp,
    fixtures tests/golden/pria_ref_large.json): the plain-PRIA instance
    (N = 50, T = 200) and two odd-T shapes of the batched mixture (37 x 151,
    100 x 299), whose z slices hold up to 29,900 entries and take the y/z
    kernel's memory path with its LDS cache.
  * PRAQ against the oracle at the mixture's sizes (tests/golden/praq_c3.json:
    n = 2e4 and 1e5), whose bracket must also hold the max-flow roof dual.
  * The reference's special cases inside real solves: a NaN gap keeps
    iterating to max_iter (qubo/praq-jax/tests/test_structure.py:191-198),
    and op_steps <= 0 takes the size rule's check interval
    (qubo/praq-jax/tests/test_solver.py:138-151).

Parity bar (FP64 iterates): iterations equal to the reference solver's (PRIA)
or the oracle's (PRAQ) on the plain device tier and in the batched mode;
objectives within 1e-9 relative; the solution checksums (sum, sum of squares,
dot product with fixed weights) within 1e-8 per entry on average (sum of
squares 2e-8), the elementwise bar of the small-instance tests.  FP32 iterates
with FP64 reductions: status 0, gap <= 1e-6, and the bracket meets the
reference's bracket to 2e-6 relative (both hold the LP optimum; FP32 bounds
carry the iterates' rounding).  An instance's batched and
served results are bitwise the same alone and in the mix."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from test_gpu_batch import _same_bits

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def _weights(shape):
    k = np.arange(int(np.prod(shape)), dtype=np.float64)
    return np.sin(0.7 * k + 0.3).reshape(shape)


def _check_sums(a, expected, what):
    """expected = [sum, sum of squares, dot with the fixed weights]."""
    a = np.asarray(a, dtype=np.float64)
    count = a.size
    got = [float(a.sum()), float((a * a).sum()), float((a * _weights(a.shape)).sum())]
    for q, bound in enumerate((1e-8, 2e-8, 1e-8)):
        assert abs(got[q] - expected[q]) <= bound * count, (what, q, got[q], expected[q])


def _pria_cases():
    from reformulation_amd import workloads
    for row in json.loads((GOLDEN / "pria_ref_large.json").read_text()):
        yield row, workloads.pria_c1(row["seed"], row["items"], row["periods"])


def _praq_cases():
    from reformulation_amd import workloads
    for row in json.loads((GOLDEN / "praq_c3.json").read_text()):
        yield row, workloads.praq_er(row["seed"], row["n"], row["m"])


def _check_pria_fp64(info, x, y, z, row):
    assert info.status == 0
    assert info.iterations == row["iterations"], (row["items"], row["periods"], info.iterations)
    sc = 1.0 + abs(row["primal"])
    assert abs(info.upper_iterate - row["primal"]) <= 1e-9 * sc
    assert abs(info.lower - row["dual"]) <= 1e-9 * sc
    _check_sums(x, row["x"], "x")
    _check_sums(y, row["y"], "y")
    _check_sums(z, row["z"], "z")


def _check_praq_fp64(info, x, row):
    assert info.status == 0
    assert info.iterations == row["iterations"] and info.restarts == row["restarts"]
    sc = 1.0 + abs(row["upper"])
    assert abs(info.upper - row["upper"]) <= 1e-9 * sc
    assert abs(info.lower - row["lower"]) <= 1e-9 * sc
    _check_sums(x, row["x"], "x")
    assert info.lower <= row["roof_dual"] + 1e-9 * sc <= info.upper + 2e-9 * sc


def _brackets_meet(lower, upper, ref_lower, ref_upper):
    """Both brackets hold the LP optimum, so they meet; FP32-iterate bounds
    carry the iterates' rounding (the FP32 bar, 2e-6 relative)."""
    sc = 1.0 + abs(ref_upper)
    assert lower <= ref_upper + 2e-6 * sc and ref_lower <= upper + 2e-6 * sc


def test_plain_pria_fp64_matches_the_reference_solver_at_workload_sizes():
    from reformulation_amd import pria
    for row, inst in _pria_cases():
        print(f"plain pria {row['items']}x{row['periods']}", flush=True)
        params = pria.Parameters.with_auto_step(None, 1e-6, tier=pria.TIER_DEVICE,
                                                time_limit=120.0)
        x, y, z, info = pria.solve_with_info(inst.problem(), params)
        assert info.tier == pria.TIER_DEVICE
        _check_pria_fp64(info, x, y, z, row)


def test_plain_praq_fp64_matches_the_oracle_at_workload_sizes():
    from reformulation_amd import praq
    for row, inst in _praq_cases():
        print(f"plain praq n={row['n']}", flush=True)
        params = praq.Parameters.with_auto_step(None, 1e-6, time_limit=120.0)
        x, _, info = praq.solve_with_info(inst.problem(), params)
        _check_praq_fp64(info, x, row)


def _scale_batch(f32, only=None):
    from reformulation_amd import praq, pria
    from reformulation_amd.batch import Batch
    b = Batch(use_float32=f32)
    index = []
    for row, inst in _pria_cases():
        if only is None or ("pria", row["seed"]) in only:
            index.append(("pria", row, b.add_pria(
                inst.problem(), pria.Parameters.with_auto_step(100_000, 1e-6))))
    for row, inst in _praq_cases():
        if only is None or ("praq", row["seed"]) in only:
            index.append(("praq", row, b.add_praq(
                inst.problem(), praq.Parameters.with_auto_step(100_000, 1e-6))))
    b.upload()
    return b, index


def _scale_results(b, index):
    out = []
    for kind, row, q in index:
        if kind == "pria":
            x, y, z, info = b.pria_result(q)
            out.append((x, (y, z), info))
        else:
            x, y, info = b.praq_result(q)
            out.append((x, y, info))
    return out


def test_batched_fp64_matches_the_references_at_workload_sizes():
    b, index = _scale_batch(False)
    print("batched fp64 at workload sizes", flush=True)
    b.solve("batched")
    for (kind, row, _), (x, dual, info) in zip(index, _scale_results(b, index)):
        if kind == "pria":
            _check_pria_fp64(info, x, dual[0], dual[1], row)
        else:
            _check_praq_fp64(info, x, row)
    b.close()


def test_batched_and_served_fp32_at_workload_sizes():
    """FP32 iterates with the FP64 handover: the brackets meet the
    references'; each result is bitwise the same alone, in the mix and
    served."""
    from reformulation_amd import praq, pria
    from reformulation_amd.serve import Server
    b, index = _scale_batch(True)
    print("batched fp32 at workload sizes", flush=True)
    b.solve("batched")
    mixed = _scale_results(b, index)
    b.close()
    for (kind, row, _), (_, _, info) in zip(index, mixed):
        assert info.status == 0 and info.relative_gap <= 1e-6, (kind, row["seed"])
        if kind == "pria":
            _brackets_meet(info.lower, info.upper, row["dual"], row["primal"])
        else:
            _brackets_meet(info.lower, info.upper, row["lower"], row["upper"])
            assert info.lower <= row["roof_dual"] + 2e-6 * abs(row["roof_dual"])
    # alone: the largest PRIA member (long z slices) and the largest PRAQ one
    for pick in ([2], [4]):
        kind, row, _ = index[pick[0]]
        b2, idx2 = _scale_batch(True, only={(kind, row["seed"])})
        b2.solve("batched")
        alone = _scale_results(b2, idx2)[0]
        b2.close()
        assert _same_bits(alone[0], mixed[pick[0]][0]), pick
        assert _same_bits(alone[1], mixed[pick[0]][1]), pick
        assert alone[2].iterations == mixed[pick[0]][2].iterations
    # served, all at once on a small server
    print("served fp32 at workload sizes", flush=True)
    with Server(use_float32=True, host_threads=2) as srv:
        tickets = []
        for kind, row, _ in index:
            if kind == "pria":
                inst = next(i for r, i in _pria_cases() if r["seed"] == row["seed"])
                tickets.append(srv.submit_pria(inst.problem(),
                                               pria.Parameters.with_auto_step(100_000, 1e-6)))
            else:
                inst = next(i for r, i in _praq_cases() if r["seed"] == row["seed"])
                tickets.append(srv.submit_praq(inst.problem(),
                                               praq.Parameters.with_auto_step(100_000, 1e-6)))
        srv.wait_idle(120.0)
        for (kind, row, _), t, (bx, bdual, binfo) in zip(index, tickets, mixed):
            if kind == "pria":
                x, y, z, info = srv.pria_result(t)
                dual = (y, z)
            else:
                x, dual, info = srv.praq_result(t)
            assert info.status == 0, (kind, row["seed"])
            assert _same_bits(x, bx) and _same_bits(dual, bdual), (kind, row["seed"])
            assert info.iterations == binfo.iterations
            assert info.lower == binfo.lower and info.upper == binfo.upper


def _nan_praq():
    # the reference's instance: W = [[nan, 1], [1, -1]]
    return 2, np.array([np.nan, -1.0]), np.array([0]), np.array([1]), np.array([1.0])


def _nan_pria():
    from helpers import small_jrp
    k, k0, h = small_jrp(21, 3, 6)
    h = h.copy()
    h[1, 0, 4] = np.nan
    return k, k0, h


@pytest.mark.parametrize("tier", [1, 2])
def test_nan_gap_keeps_iterating_to_max_iter(tier):
    """A NaN gap is not convergence: ~(gap <= tol) holds, so the solve runs
    to max_iter (64) and reports the iteration limit."""
    from reformulation_amd import praq, pria
    n, d, r, c, v = _nan_praq()
    x, _, info = praq.solve_with_info(
        praq.QuboProblem.from_sparse(n, d, r, c, v),
        praq.Parameters.with_auto_step(64, 1e-9, tier=tier, time_limit=30.0))
    assert info.iterations >= 64 and info.status == 1, (info.iterations, info.status)
    k, k0, h = _nan_pria()
    _, _, _, pinfo = pria.solve_with_info(
        pria.JrpProblem.new(k, k0, h),
        pria.Parameters.with_auto_step(64, 1e-9, tier=tier, time_limit=30.0))
    assert pinfo.iterations == 64 and pinfo.status == 1, (pinfo.iterations, pinfo.status)


@pytest.mark.parametrize("f32", [False, True])
def test_nan_gap_in_the_batched_mode(f32):
    from reformulation_amd import praq, pria
    from reformulation_amd.batch import Batch
    from helpers import small_jrp
    b = Batch(use_float32=f32)
    qn = b.add_praq(praq.QuboProblem.from_sparse(*_nan_praq()),
                    praq.Parameters.with_auto_step(64, 1e-9))
    pn = b.add_pria(pria.JrpProblem.new(*_nan_pria()), pria.Parameters.with_auto_step(64, 1e-9))
    k, k0, h = small_jrp(22, 4, 9)
    pg = b.add_pria(pria.JrpProblem.new(k, k0, h), pria.Parameters.with_auto_step(None, 1e-8))
    b.upload()
    b.solve("batched")
    _, _, info = b.praq_result(qn)
    assert info.iterations >= 64 and info.status == 1
    _, _, _, info = b.pria_result(pn)
    assert info.iterations == 64 and info.status == 1
    _, _, _, good = b.pria_result(pg)
    o = oracle.pria_solve(k, k0, h, oracle.pria_params(tolerance=1e-8))
    assert good.status == 0
    if not f32:
        assert good.iterations == o.iterations
    b.close()


def test_op_steps_zero_takes_the_size_rule_in_real_solves():
    """op_steps <= 0: the check interval of the size rule, 2 at n = 60 and 3
    at n = 1e5 (round(3 log2(2)) = 3), on both plain tiers and in the
    batched mode, with the oracle's iterations."""
    from reformulation_amd import praq, workloads
    from reformulation_amd.batch import Batch
    from helpers import random_qubo
    cases = [random_qubo(3, 60, 0.2), workloads.praq_er(410, 100_000, 400_000)]
    cases[1] = (cases[1].n, cases[1].diag, cases[1].rows, cases[1].cols, cases[1].vals)
    assert praq.check_interval(60) == 2 and praq.check_interval(100_000) == 3
    b = Batch(use_float32=False)
    index = []
    expect = []
    for inst in cases:
        o = oracle.praq_solve(*inst, params=oracle.praq_params(tolerance=1e-7, op_steps=0))
        interval = praq.check_interval(inst[0])
        assert o.status == 0 and o.iterations > 0 and o.iterations % interval == 0
        expect.append(o)
        problem = praq.QuboProblem.from_sparse(*inst)
        for tier in (1, 2):
            print(f"op_steps 0: n={inst[0]} tier={tier}", flush=True)
            _, _, info = praq.solve_with_info(
                problem, praq.Parameters.with_auto_step(None, 1e-7, op_steps=0, tier=tier,
                                                        time_limit=60.0))
            assert info.iterations == o.iterations and info.restarts == o.restarts
            assert abs(info.upper - o.upper) <= 1e-9 * (1 + abs(o.upper))
        index.append(b.add_praq(problem, praq.Parameters.with_auto_step(None, 1e-7,
                                                                        op_steps=0)))
    b.upload()
    b.solve("batched")
    for q, o in zip(index, expect):
        _, _, info = b.praq_result(q)
        assert info.iterations == o.iterations and info.restarts == o.restarts
        assert abs(info.upper - o.upper) <= 1e-9 * (1 + abs(o.upper))
    b.close()


@pytest.mark.parametrize("shape", [(2, 700), (3, 701)])
def test_periods_beyond_640_match_the_oracle(shape):
    """T' above 64 VR = 640: y rows longer than a lane group's registers go
    to the wavefront-per-row memory path.  FP64 on the device tier and in the
    batched mode: iterations equal the oracle's, bounds to 1e-9, x to 1e-8;
    FP32 iterates in the batched mode close the bracket on the oracle's.  The
    one-workgroup tier, limited to T' <= 640, refuses the instance."""
    import dataclasses
    from reformulation_amd import pria, workloads
    from reformulation_amd.batch import Batch
    n, t = shape
    j = workloads.pria_c1(500, n, t)
    problem = j.problem()
    print(f"long periods {n}x{t}: oracle", flush=True)
    o = oracle.pria_solve(j.order_costs, j.joint_order_cost, j.holding_costs,
                          oracle.pria_params(tolerance=1e-6))
    sc = 1.0 + abs(o.upper)
    params = pria.Parameters.with_auto_step(50_000, 1e-6)
    x, _, _, info = pria.solve_with_info(problem, dataclasses.replace(params, tier=2))
    assert info.status == 0 and info.tier == 2
    assert info.iterations == o.iterations
    assert abs(info.upper - o.upper) <= 1e-9 * sc and abs(info.lower - o.lower) <= 1e-9 * sc
    assert np.max(np.abs(x - o.x)) <= 1e-8
    with pytest.raises(ValueError):
        pria.solve_with_info(problem, dataclasses.replace(params, tier=1))
    for f32 in (False, True):
        print(f"long periods {n}x{t}: batched f32={f32}", flush=True)
        b = Batch(use_float32=f32)
        q = b.add_pria(problem, params)
        b.upload()
        b.solve("batched")
        _, _, _, bi = b.pria_result(q, with_solution=False)
        b.close()
        assert bi.status == 0
        if f32:
            assert bi.lower <= o.upper + 2e-6 * sc and bi.upper >= o.lower - 2e-6 * sc
        else:
            assert bi.iterations == o.iterations
            assert abs(bi.upper - o.upper) <= 1e-9 * sc


@pytest.mark.parametrize("joint", ["k0", "zero"])
def test_device_tier_chunked_z_pass_agrees_with_the_batched_tier(joint):
    """N x T = 60 x 380 (ND = 4.3M, slices of up to 22,800 entries): the
    device tier runs its z projection over chunks of 4096 entries with
    multi-block Michelot rounds (one workgroup finishing any slice not
    settled after four), the batched tier one workgroup per slice.  FP64
    iterations, restarts and bounds agree; with K0 = 0 the z part is 0."""
    from reformulation_amd import pria, workloads
    from reformulation_amd.batch import Batch
    inst = workloads.pria_c1(7, 60, 380)
    k0 = inst.joint_order_cost if joint == "k0" else 0.0
    problem = pria.JrpProblem.new(inst.order_costs, k0, inst.holding_costs)
    print(f"device tier 60x380 K0={k0}", flush=True)
    params = pria.Parameters.with_auto_step(None, 1e-6, tier=pria.TIER_DEVICE, time_limit=120.0)
    x, y, z, info = pria.solve_with_info(problem, params)
    assert info.tier == pria.TIER_DEVICE and info.status == 0
    print("batched 60x380", flush=True)
    b = Batch(use_float32=False)
    q = b.add_pria(problem, pria.Parameters.with_auto_step(None, 1e-6))
    b.upload()
    b.solve("batched")
    bx, by, bz, binfo = b.pria_result(q)
    b.close()
    assert binfo.status == 0
    assert info.iterations == binfo.iterations and info.restarts == binfo.restarts
    sc = 1.0 + abs(binfo.upper)
    assert abs(info.upper - binfo.upper) <= 1e-9 * sc
    assert abs(info.lower - binfo.lower) <= 1e-9 * sc
    assert np.max(np.abs(x - bx)) <= 1e-7
    if joint == "zero":
        assert not np.any(z) and not np.any(bz)
