"""The batched mode (block tables over a ragged batch, per-instance Halpern
state on the device, hipGraph rounds) against the plain solver and the
oracle, through the batch C ABI.  Marked gpu.

Parity bar: FP64 batched iterations equal the oracle's and the plain
solver's; objectives within 1e-9 relative of the oracle's; an instance's
batched result is bitwise the same alone and inside a mix, and the same on a
repeated solve.  FP32 iterates: the bracket closes to the requested gap and
the objective agrees with the plain FP32 solver's to 2e-6 relative."""
import numpy as np
import pytest

import oracle
from helpers import random_qubo, small_jrp

pytestmark = pytest.mark.gpu


def _members():
    """A ragged mix: PRAQ instances of the mixture's kind at small sizes and
    tiny enumeration-size ones, PRIA instances with odd and even T."""
    from reformulation_amd import workloads
    out = []
    for seed, (n, m) in enumerate([(1200, 4000), (3000, 9000), (800, 6400)]):
        w = workloads.praq_er(100 + seed, n, m)
        out.append(("praq", (w.n, w.diag, w.rows, w.cols, w.vals)))
    out.append(("praq", random_qubo(7, 6, 0.7)))
    for seed, (n, t) in enumerate([(12, 41), (20, 60), (8, 17), (30, 90)]):
        j = workloads.pria_c1(200 + seed, n, t)
        out.append(("pria", (j.order_costs, j.joint_order_cost, j.holding_costs)))
    out.append(("pria", small_jrp(3, 5, 9)))
    return out


def _batch(members, f32, tol=1e-6):
    b, index = _batch_open(members, f32, tol)
    b.upload()
    return b, index


def _batch_open(members, f32, tol=1e-6):
    """The batch with the members added, not yet uploaded."""
    from reformulation_amd import praq, pria
    from reformulation_amd.batch import Batch
    b = Batch(use_float32=f32)
    index = []
    for kind, inst in members:
        if kind == "praq":
            n, diag, rows, cols, vals = inst
            p = praq.QuboProblem.from_sparse(n, diag, rows, cols, vals)
            index.append(("praq", b.add_praq(p, praq.Parameters.with_auto_step(50_000, tol))))
        else:
            k, k0, h = inst
            p = pria.JrpProblem.new(k, k0, h)
            index.append(("pria", b.add_pria(p, pria.Parameters.with_auto_step(50_000, tol))))
    return b, index


def _results(b, index):
    out = []
    for kind, q in index:
        if kind == "praq":
            x, y, info = b.praq_result(q)
            out.append((x, y, info))
        else:
            x, y, z, info = b.pria_result(q)
            out.append((x, (y, z), info))
    return out


def _same_bits(a, b):
    if isinstance(a, tuple):
        return all(_same_bits(u, v) for u, v in zip(a, b))
    return a is None and b is None or np.array_equal(a, b)


@pytest.mark.parametrize("f32", [False, True])
def test_batched_equals_plain_solver(f32):
    members = _members()
    b, index = _batch(members, f32)
    b.solve("batched")
    batched = _results(b, index)
    b.solve("sequential")
    plain = _results(b, index)
    for (kind, _), (xb, _, ib), (xs, _, is_) in zip(index, batched, plain):
        assert ib.status == 0 and is_.status == 0
        assert ib.relative_gap <= 1e-6
        sc = 1.0 + abs(is_.upper)
        if f32:
            # FP32 iterates: the plain tiers and the batched kernels reduce in
            # different orders, so the trajectories part at FP32 rounding; both
            # brackets close to 1e-6 around the same LP value
            assert abs(ib.upper - is_.upper) <= 2e-6 * sc
            continue
        assert ib.iterations == is_.iterations, kind
        assert ib.restarts == is_.restarts, kind
        assert abs(ib.upper - is_.upper) <= 1e-9 * sc
        assert abs(ib.lower - is_.lower) <= 1e-9 * sc
        assert np.max(np.abs(xb - xs)) <= 1e-8
    b.close()


def test_batched_fp64_matches_the_oracle():
    members = _members()
    b, index = _batch(members, False)
    b.solve("batched")
    for (kind, inst), (x, _, info) in zip(members, _results(b, index)):
        if kind == "praq":
            o = oracle.praq_solve(*inst, params=oracle.praq_params(tolerance=1e-6))
        else:
            k, k0, h = inst
            o = oracle.pria_solve(k, k0, h, oracle.pria_params(tolerance=1e-6))
        assert info.iterations == o.iterations and info.restarts == o.restarts, kind
        sc = 1.0 + abs(o.upper)
        assert abs(info.upper - o.upper) <= 1e-9 * sc, kind
        assert abs(info.lower - o.lower) <= 1e-9 * sc, kind
        assert np.max(np.abs(x - o.x)) <= 1e-7, kind
    b.close()


@pytest.mark.parametrize("f32", [False, True])
def test_result_is_the_same_alone_in_the_mix_and_on_a_second_solve(f32):
    members = _members()
    b, index = _batch(members, f32)
    b.solve("batched")
    mixed = _results(b, index)
    b.solve("batched")
    again = _results(b, index)
    for u, v in zip(mixed, again):
        assert _same_bits(u[0], v[0]) and _same_bits(u[1], v[1])
        assert u[2].iterations == v[2].iterations
    for pick in ([0], [4, 1], [5, 8, 3]):
        sub = [members[q] for q in pick]
        b2, idx2 = _batch(sub, f32)
        b2.solve("batched")
        for q, (x, dual, info) in zip(pick, _results(b2, idx2)):
            assert _same_bits(x, mixed[q][0]) and _same_bits(dual, mixed[q][1]), q
            assert info.iterations == mixed[q][2].iterations
            assert info.upper == mixed[q][2].upper and info.lower == mixed[q][2].lower
        b2.close()
    b.close()


@pytest.mark.parametrize("f32", [False, True])
def test_a_long_horizon_member_changes_no_other_result(f32):
    """A member with T = 331 makes the batch take the wide y/z register
    shape, FP32 x tiles without registers held and FP64 x tiles as two
    256-thread half-tile workgroups (a 32-column tile no longer fits two to
    a CU's LDS), all with the same sums: the other members' results are
    bitwise those they have without it."""
    from reformulation_amd import workloads
    small = [m for m in _members() if m[0] == "pria"][:3]
    j = workloads.pria_c1(9, 6, 331)
    big = ("pria", (j.order_costs, j.joint_order_cost, j.holding_costs))
    b1, i1 = _batch(small, f32)
    b1.solve("batched")
    alone = _results(b1, i1)
    b1.close()
    b2, i2 = _batch(small + [big], f32)
    b2.solve("batched")
    mixed = _results(b2, i2)
    b2.close()
    assert mixed[-1][2].status == 0
    for u, v in zip(alone, mixed):
        assert _same_bits(u[0], v[0]) and _same_bits(u[1], v[1])
        assert u[2].iterations == v[2].iterations
        assert u[2].upper == v[2].upper and u[2].lower == v[2].lower


def test_timed_launches_give_the_same_results_and_count_every_step():
    members = _members()
    b, index = _batch(members, True)
    b.solve("batched")
    graph = _results(b, index)
    b.set_timing(True)
    b.solve("batched")
    timed = _results(b, index)
    stats = b.kernel_stats()
    b.set_timing(False)
    for u, v in zip(graph, timed):
        assert _same_bits(u[0], v[0]) and u[2].iterations == v[2].iterations
    names = {s["name"] for s in stats}
    assert {"wpraq_node", "wpraq_edge", "wpria_x", "wpria_yz"} <= names
    for s in stats:
        assert s["launches"] > 0 and s["ms"] > 0.0 and s["bytes"] > 0.0
    b.close()


def _edge_members():
    """The special cases of the reference's tests, mixed in one batch: one
    node (d = -3: x = 1), isolated nodes, the zero instance, a max-cut form
    (ends at iteration 0), no couplings at all; PRIA with K0 = 0 (z = 0),
    one item, one period (padded to two), odd T."""
    n, _, rows, cols, vals = random_qubo(7, 9, 0.7)
    vals = np.abs(vals)
    diag = np.zeros(n)
    np.add.at(diag, rows, -vals)
    np.add.at(diag, cols, -vals)
    empty = np.zeros(0, np.int64)
    out = [("praq", (1, np.array([-3.0]), empty, empty, np.zeros(0))),
           ("praq", (5, np.array([2.0, -1.0, 0.0, 3.0, -4.0]), empty, empty, np.zeros(0))),
           ("praq", (4, np.zeros(4), np.array([0, 1]), np.array([2, 3]), np.zeros(2))),
           ("praq", (n, diag, rows, cols, vals)),
           ("praq", random_qubo(9, 6, 0.5))]
    k, _, h = small_jrp(4, 3, 8)
    out.append(("pria", (k, 0.0, h)))
    out.append(("pria", small_jrp(5, 1, 6)))
    out.append(("pria", small_jrp(6, 3, 1)))
    out.append(("pria", small_jrp(7, 2, 5)))
    return out


@pytest.mark.parametrize("f32", [False, True])
def test_special_cases_batched_and_served_match_the_oracle(f32):
    from reformulation_amd.serve import Server
    from test_gpu_serve import _problems, _served, _submit
    members = _edge_members()
    b, index = _batch(members, f32, tol=1e-8)
    b.solve("batched")
    batched = _results(b, index)
    b.close()
    with Server(use_float32=f32, host_threads=2) as srv:
        tickets = [_submit(srv, item) for item in _problems(members, tol=1e-8)]
        srv.wait_idle(60.0)
        served = [_served(srv, kind, t) for (kind, _), t in zip(members, tickets)]
    for q, ((kind, inst), (x, dual, info)) in enumerate(zip(members, batched)):
        if kind == "praq":
            o = oracle.praq_solve(*inst, params=oracle.praq_params(tolerance=1e-8))
        else:
            k, k0, h = inst
            o = oracle.pria_solve(k, k0, h, oracle.pria_params(tolerance=1e-8))
        assert info.status == 0, q
        sc = 1.0 + abs(o.upper)
        if not f32:
            assert info.iterations == o.iterations, (q, info.iterations, o.iterations)
            assert abs(info.upper - o.upper) <= 1e-9 * sc, q
        else:
            assert abs(info.upper - o.upper) <= 2e-6 * sc, q
        if kind == "pria" and inst[1] == 0.0:
            assert np.all(dual[1] == 0.0), q          # z = 0 when K0 = 0
        sx, sdual, sinfo = served[q]
        assert _same_bits(sx, x) and _same_bits(sdual, dual), q
        assert sinfo.iterations == info.iterations, q
    assert abs(batched[0][2].upper + 3.0) < 1e-8       # n = 1, d = -3
    if not f32:
        # the max-cut form certifies at its first check; FP32 iterates round
        # its bounds apart by more than 1e-8 and finish after the handover
        assert batched[3][2].iterations == 0


def _limited():
    """A PRIA instance that cannot reach its tolerance (1e-15) within its
    time limit (0.05 s): N = 50, T = 200."""
    from reformulation_amd import pria, workloads
    j = workloads.pria_c1(900, 50, 200)
    params = pria.Parameters.with_auto_step(100_000_000, 1e-15, time_limit=0.05)
    return pria.JrpProblem.new(j.order_costs, j.joint_order_cost, j.holding_costs), params


@pytest.mark.parametrize("f32", [False, True])
@pytest.mark.parametrize("mode", ["batched", "sequential"])
def test_instance_time_limit_stops_only_that_instance(f32, mode):
    """An instance's own time limit (options.time_limit) ends it at its current
    point with status 2; the other instances of the batch run to their
    tolerance with the results they have without it."""
    members = _members()
    b0, index0 = _batch(members, f32)
    b0.solve(mode)
    reference = _results(b0, index0)
    b0.close()
    b, index = _batch_open(members, f32)
    problem, params = _limited()
    q_limited = b.add_pria(problem, params)
    b.upload()
    print(f"time limit: mode={mode} f32={f32}", flush=True)
    b.solve(mode)
    for q, (u, v) in enumerate(zip(_results(b, index), reference)):
        assert u[2].status == 0, q
        assert _same_bits(u[0], v[0]) and _same_bits(u[1], v[1]), q
        assert u[2].iterations == v[2].iterations, q
    x, y, z, info = b.pria_result(q_limited)
    assert info.status == 2
    assert 0 < info.iterations < params.max_iter
    assert np.all(np.isfinite(x)) and np.all(np.isfinite(y)) and np.all(np.isfinite(z))
    assert np.isfinite(info.lower) and np.isfinite(info.upper)
    assert info.relative_gap > 1e-15
    b.close()


@pytest.mark.parametrize("f32", [False, True])
def test_compaction_changes_no_result(f32):
    """Tables rebuilt over the running instances as others finish give every
    instance bitwise the answer, iterations and bounds of the uncompacted
    solve, and the same algorithmic bytes per step kernel."""
    members = _members() + _members()[:4]
    b, index = _batch(members, f32)
    b.set_timing(True)
    b.set_compaction(False)
    b.solve("batched")
    plain = _results(b, index)
    plain_stats = {s["name"]: s for s in b.kernel_stats()}
    b.set_compaction(True)
    b.solve("batched")
    compact = _results(b, index)
    stats = {s["name"]: s for s in b.kernel_stats()}
    for q, (u, v) in enumerate(zip(compact, plain)):
        assert _same_bits(u[0], v[0]) and _same_bits(u[1], v[1]), q
        assert u[2].iterations == v[2].iterations, q
        assert u[2].lower == v[2].lower and u[2].upper == v[2].upper, q
    for name, s in stats.items():
        assert s["bytes"] == plain_stats[name]["bytes"], name
    b.close()
