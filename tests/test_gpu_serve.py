"""The serving loop (include/reform_serve.h) against the batched mode, through
the server C ABI.  Marked gpu.

Parity bar: a request served by the server -- admitted while others are in
flight, its round shared with whatever else is resident, retired and
backfilled -- returns bitwise the answer, iteration count and bounds the
batched mode gives the same instance, in FP64 and with FP32 iterates: an
instance's arithmetic does not depend on what else is on the device."""
import time

import numpy as np
import pytest

from test_gpu_batch import _members, _batch, _results, _same_bits

pytestmark = pytest.mark.gpu


def _problems(members, tol=1e-6):
    from reformulation_amd import praq, pria
    out = []
    for kind, inst in members:
        if kind == "praq":
            n, diag, rows, cols, vals = inst
            p = praq.QuboProblem.from_sparse(n, diag, rows, cols, vals)
            out.append(("praq", p, praq.Parameters.with_auto_step(50_000, tol)))
        else:
            k, k0, h = inst
            out.append(("pria", pria.JrpProblem.new(k, k0, h),
                        pria.Parameters.with_auto_step(50_000, tol)))
    return out


def _submit(srv, item):
    kind, problem, params = item
    return srv.submit_praq(problem, params) if kind == "praq" else srv.submit_pria(problem, params)


def _served(srv, kind, t):
    if kind == "praq":
        x, y, info = srv.praq_result(t)
        return x, y, info
    x, y, z, info = srv.pria_result(t)
    return x, (y, z), info


def _check_same(served, batched, q):
    x, dual, info = served
    bx, bdual, binfo = batched
    assert info.status == 0, q
    assert _same_bits(x, bx), q
    assert _same_bits(dual, bdual), q
    assert info.iterations == binfo.iterations, q
    assert info.lower == binfo.lower and info.upper == binfo.upper, q


@pytest.mark.parametrize("f32", [False, True])
def test_served_equals_batched_with_staggered_arrivals(f32):
    from reformulation_amd.serve import Server
    members = _members()
    b, index = _batch(members, f32)
    b.solve("batched")
    reference = _results(b, index)
    b.close()
    items = _problems(members)
    print(f"serve f32={f32}: {len(items)} requests, staggered", flush=True)
    with Server(use_float32=f32, host_threads=2) as srv:
        tickets = []
        for q, item in enumerate(items):
            tickets.append(_submit(srv, item))
            time.sleep(0.003 * (q % 3))   # later requests join rounds in progress
        srv.wait_idle(120.0)
        stats = srv.stats()
        for q, ((kind, _, _), t) in enumerate(zip(items, tickets)):
            _check_same(_served(srv, kind, t), reference[q], q)
            srv.release(t)
        assert stats["completed"] == len(items) and stats["rejected"] == 0
        assert stats["bytes_in_use"] == 0 and stats["in_flight"] == 0


def test_backfill_under_a_small_budget():
    """A budget that holds about two of the instances: requests wait for
    memory, are admitted as others retire, and still return the batched
    answers; the pool never exceeds its budget."""
    from reformulation_amd.serve import Server
    members = _members()
    b, index = _batch(members, True)
    b.solve("batched")
    reference = _results(b, index)
    b.close()
    items = _problems(members)
    with Server(use_float32=True, host_threads=2) as srv:
        tickets = [_submit(srv, item) for item in items]
        srv.wait_idle(120.0)
        sizes = [srv.info(t).device_bytes for t in tickets]
    budget = 2 * max(sizes) + (1 << 20)
    print(f"backfill: budget {budget} bytes for requests of {sizes}", flush=True)
    with Server(use_float32=True, host_threads=2, hbm_budget=budget, max_in_flight=3) as srv:
        tickets = [_submit(srv, item) for item in items]
        srv.wait_idle(120.0)
        infos = [srv.info(t) for t in tickets]
        stats = srv.stats()
        assert stats["bytes_peak"] <= stats["bytes_budget"] <= budget
        assert stats["completed"] == len(items)
        for q, ((kind, _, _), t) in enumerate(zip(items, tickets)):
            _check_same(_served(srv, kind, t), reference[q], q)
        # at most three resident at once: some request was admitted only after
        # an earlier one had completed
        admitted = np.array([i.admitted for i in infos])
        completed = np.array([i.completed for i in infos])
        assert any(admitted[q] >= completed[:q].min() for q in range(3, len(infos)))
        for q in range(len(infos)):
            live = np.sum((admitted <= admitted[q]) & (completed > admitted[q]))
            assert live <= 3, q


def test_a_request_larger_than_the_budget_is_rejected_and_others_proceed():
    from reformulation_amd import praq, workloads
    from reformulation_amd.serve import Server
    big = workloads.praq_er(5, 20000, 80000)
    small = workloads.praq_er(6, 500, 1500)
    params = praq.Parameters.with_auto_step(50_000, 1e-6)
    with Server(use_float32=False, host_threads=1, hbm_budget=4 << 20) as srv:
        t_big = srv.submit_praq(big.problem(), params)
        t_small = srv.submit_praq(small.problem(), params)
        srv.wait_idle(60.0)
        assert srv.info(t_big).status == -2
        assert srv.info(t_small).status == 0
        with pytest.raises(RuntimeError):
            srv.praq_result(t_big)
        assert srv.stats()["rejected"] == 1


def test_open_loop_driver_and_borrowed_inputs():
    """Poisson arrivals through open_loop with borrowed inputs: every request
    completes, latencies are positive, results match the batched mode."""
    from reformulation_amd import serve
    members = _members()
    b, index = _batch(members, True)
    b.solve("batched")
    reference = _results(b, index)
    b.close()
    items = _problems(members)
    arrivals = serve.poisson_arrivals(400.0, 0.1, seed=3)
    arrivals = arrivals[: 3 * len(items)]
    seen = {}
    with serve.Server(use_float32=True, host_threads=2, borrow_inputs=True) as srv:
        def submit(q):
            return _submit(srv, items[q % len(items)])

        def on_done(q, t):
            seen[q] = _served(srv, items[q % len(items)][0], t)

        arrive, done, status = serve.open_loop(srv, arrivals, submit, 120.0, on_done)
    assert np.all(status == 0)
    assert np.all(done > arrive)
    s = serve.summarize(arrive, done)
    assert s["requests"] == len(arrivals) and s["p99_ms"] >= s["p50_ms"] > 0
    for q, res in seen.items():
        _check_same(res, reference[q % len(items)], q)
    assert len(seen) == len(arrivals)


def test_request_time_limit_stops_only_that_request():
    """A request's own time limit counts from its admission: past it the
    request finishes with status 2 at its current point, between two rounds;
    the requests beside it return the batched answers."""
    from reformulation_amd.serve import Server
    from test_gpu_batch import _limited
    members = _members()
    b, index = _batch(members, True)
    b.solve("batched")
    reference = _results(b, index)
    b.close()
    items = _problems(members)
    problem, params = _limited()
    with Server(use_float32=True, host_threads=2) as srv:
        t_limited = srv.submit_pria(problem, params)
        tickets = [_submit(srv, item) for item in items]
        srv.wait_idle(120.0)
        for q, ((kind, _, _), t) in enumerate(zip(items, tickets)):
            _check_same(_served(srv, kind, t), reference[q], q)
        info = srv.info(t_limited)
        assert info.status == 2
        assert 0 < info.iterations < params.max_iter
        assert info.completed - info.admitted >= params.time_limit
        x, y, z, rinfo = srv.pria_result(t_limited)
        assert np.all(np.isfinite(x)) and np.all(np.isfinite(z))
        assert np.isfinite(rinfo.lower) and np.isfinite(rinfo.upper)


def test_submit_rejects_invalid_parameters():
    """The parameters praq_solve_ex / pria_solve_ex reject are rejected at
    submission with an argument error (-1), before any admission."""
    from dataclasses import replace
    from reformulation_amd import praq, pria
    from reformulation_amd.serve import Server
    members = _problems(_members()[:1] + _members()[4:5])
    (_, qp, qparams), (_, rp, rparams) = members
    with Server(use_float32=True, host_threads=1) as srv:
        for bad in (replace(qparams, omega=-1.0), replace(qparams, tolerance=float("nan")),
                    replace(qparams, theta=float("inf"))):
            with pytest.raises(ValueError, match=r"\(-1\)"):
                srv.submit_praq(qp, bad)
        for bad in (replace(rparams, omega=0.0), replace(rparams, tolerance=float("nan"))):
            with pytest.raises(ValueError, match=r"\(-1\)"):
                srv.submit_pria(rp, bad)
        t = srv.submit_pria(rp, rparams)
        srv.wait_idle(60.0)
        assert srv.info(t).status == 0


def test_borrowed_inputs_may_go_once_the_request_has_finished():
    """With borrowed inputs the caller's arrays need to live only until the
    request has finished: overwritten afterwards (the observable stand-in for
    freeing them), the request's info and result are unchanged."""
    from reformulation_amd.serve import Server
    items = _problems(_members()[4:6])
    with Server(use_float32=True, host_threads=1) as srv:
        tickets = [_submit(srv, item) for item in items]
        srv.wait_idle(60.0)
        before = [srv.info(t) for t in tickets]
    with Server(use_float32=True, host_threads=1, borrow_inputs=True) as srv:
        tickets = [_submit(srv, item) for item in items]
        srv.wait_idle(60.0)
        for t in tickets:
            for a in srv._keep[t]:
                a.fill(np.nan)
        for t, b in zip(tickets, before):
            info = srv.info(t)
            assert info.status == 0
            assert info.upper == b.upper and info.lower == b.lower
            assert np.isfinite(srv.pria_result(t)[3].upper)


def test_fp64_images_are_placed_at_handovers_only():
    """On an FP32 server a request takes its FP64 image at its handover: one
    that converges in FP32 (tolerance 1e-3) ends holding less than the FP64
    image alone (as much, for the tiniest), one that hands over holds both
    images (more than the FP64 image alone), the handover count equals the requests that switched, and
    the answers are still the batched ones."""
    from reformulation_amd.serve import Server
    members = _members()
    b, index = _batch(members, True)
    b.solve("batched")
    reference = _results(b, index)
    b.close()
    items = _problems(members) + _problems(members[:3] + members[4:6], tol=1e-3)
    with Server(use_float32=False, host_threads=2) as srv:
        tickets = [_submit(srv, item) for item in items]
        srv.wait_idle(120.0)
        fp64_bytes = [srv.info(t).device_bytes for t in tickets]
    print(f"lazy FP64 images: {len(items)} requests", flush=True)
    with Server(use_float32=True, host_threads=2) as srv:
        tickets = [_submit(srv, item) for item in items]
        srv.wait_idle(120.0)
        stats = srv.stats()
        infos = [srv.info(t) for t in tickets]
        for q in range(len(reference)):
            _check_same(_served(srv, items[q][0], tickets[q]), reference[q], q)
    switched = [i.switch_iteration >= 0 for i in infos]
    assert stats["handovers"] == sum(switched) > 0
    assert not all(switched)
    for q, info in enumerate(infos):
        assert info.status == 0, q
        if switched[q]:
            assert info.device_bytes > fp64_bytes[q], q
        else:   # equal for images so small that both round to the same blocks
            assert info.device_bytes <= fp64_bytes[q], q
    assert any(i.device_bytes < b for i, b, s in zip(infos, fp64_bytes, switched) if not s)
    assert stats["bytes_in_use"] == 0 and stats["in_flight"] == 0


def test_a_pool_with_room_for_one_fp64_image_at_a_time_finishes_every_request():
    """A budget that holds every request's FP32 image but only one FP64 image
    beside them: handovers wait for room between rounds, admission keeps room
    for the largest FP64 image still to come, every request finishes with the
    batched answer, and more requests are in flight at once than eager FP32 +
    FP64 images would allow."""
    from reformulation_amd.serve import Server
    members = _members()
    b, index = _batch(members, True)
    b.solve("batched")
    reference = _results(b, index)
    b.close()
    items = _problems(members)
    with Server(use_float32=False, host_threads=2) as srv:
        tickets = [_submit(srv, item) for item in items]
        srv.wait_idle(120.0)
        b64 = [srv.info(t).device_bytes for t in tickets]
    with Server(use_float32=True, host_threads=2) as srv:
        tickets = [_submit(srv, item) for item in items]
        srv.wait_idle(120.0)
        both = [srv.info(t).device_bytes for t in tickets]
        switched = [srv.info(t).switch_iteration >= 0 for t in tickets]
    b32 = [u - v if s else u for u, v, s in zip(both, b64, switched)]
    budget = sum(b32) + max(b64) + (1 << 16)
    eager = budget // max(u + v for u, v in zip(b32, b64))
    print(f"one FP64 image at a time: budget {budget}, eager fit {eager}", flush=True)
    with Server(use_float32=True, host_threads=2, hbm_budget=budget) as srv:
        tickets = [_submit(srv, item) for item in items]
        srv.wait_idle(120.0)
        stats = srv.stats()
        for q, ((kind, _, _), t) in enumerate(zip(items, tickets)):
            _check_same(_served(srv, kind, t), reference[q], q)
        assert stats["completed"] == len(items) and stats["rejected"] == 0
        assert stats["bytes_peak"] <= stats["bytes_budget"] <= budget
        assert stats["handovers"] == sum(switched)
        assert stats["in_flight_peak"] > eager


def _finished(srv, t, timeout=60.0):
    end = time.time() + timeout
    while srv.info(t).status == -1:
        assert time.time() < end, t
        time.sleep(0.005)
    return srv.info(t)


def test_cancel_before_admission_and_in_flight():
    """One state slot: a request cancelled while it waits for the slot
    finishes at once with status -6 and no answer; the one holding the slot,
    cancelled in flight, stops at its next round with status 3, its point
    as the answer and its bounds still a bracket, as at a time limit; the
    server goes on serving."""
    from reformulation_amd import pria, workloads
    from reformulation_amd.serve import Server
    j = workloads.pria_c1(5, 60, 200)
    prob = pria.JrpProblem.new(j.order_costs, j.joint_order_cost, j.holding_costs)
    # tolerance 0: runs until its bounds meet exactly, far past this test
    # (the time limit is a backstop only)
    endless = pria.Parameters.with_auto_step(10**8, 0.0, time_limit=120.0)
    quick = pria.Parameters.with_auto_step(50_000, 1e-6)
    with Server(use_float32=True, host_threads=2, max_in_flight=1) as srv:
        a = srv.submit_pria(prob, endless)
        b = srv.submit_pria(prob, quick)
        end = time.time() + 60
        while srv.info(a).admitted < 0:
            assert time.time() < end
            time.sleep(0.005)
        assert srv.cancel(b)
        ib = _finished(srv, b)
        assert ib.status == -6 and ib.admitted < 0
        with pytest.raises(RuntimeError):
            srv.pria_result(b)
        time.sleep(0.1)
        assert srv.cancel(a), srv.info(a)
        assert not srv.cancel(a)           # already cancelled
        srv.wait_idle(60)
        x, y, z, ia = srv.pria_result(a)
        assert ia.status == 3 and ia.iterations > 0
        assert all(np.isfinite(v).all() for v in (x, y, z))
        assert np.isfinite(ia.lower) and np.isfinite(ia.upper)
        assert ia.lower <= ia.upper + 1e-9 * abs(ia.upper)
        assert not srv.cancel(a) and not srv.cancel(b)   # finished
        assert not srv.cancel(10**9)                      # unknown
        c = srv.submit_pria(prob, quick)
        srv.wait_idle(60)
        assert srv.info(c).status == 0
        stats = srv.stats()
        assert stats["cancelled"] == 2 and stats["completed"] == 3


def test_cancelled_request_that_converged_first_keeps_its_answer():
    """A cancel that comes after the request has converged on the device but
    before its answer is out changes nothing: status 0 and the batched
    answer."""
    from reformulation_amd.serve import Server
    members = [m for m in _members() if m[0] == "pria"][:1]
    b, index = _batch(members, True)
    b.solve("batched")
    want = _results(b, index)[0]
    b.close()
    items = _problems(members)
    with Server(use_float32=True, host_threads=1) as srv:
        t = _submit(srv, items[0])
        srv.wait_idle(60)
        assert not srv.cancel(t)
        _check_same(_served(srv, "pria", t), want, 0)


def test_non_finite_inputs_are_rejected():
    from reformulation_amd import praq, pria
    from reformulation_amd.serve import Server
    items = _problems(_members())
    kind, p, params = items[0]
    diag = p.diag.copy()
    diag[3] = np.nan
    bad_praq = praq.QuboProblem.from_sparse(p.n(), diag, p.rows, p.cols, p.vals)
    pk = next(it for it in items if it[0] == "pria")
    h = pk[1].holding_costs.copy()
    h[0, 0, -1] = np.inf
    bad_pria = pria.JrpProblem.new(pk[1].order_costs, pk[1].joint_order_cost, h)
    with Server(use_float32=True, host_threads=2) as srv:
        t1 = srv.submit_praq(bad_praq, params)
        t2 = srv.submit_pria(bad_pria, pk[2])
        t3 = _submit(srv, pk)
        srv.wait_idle(60)
        assert srv.info(t1).status == -2 and srv.info(t2).status == -2
        assert srv.info(t3).status == 0
        assert srv.stats()["rejected"] == 2


@pytest.mark.parametrize("f32", [False, True])
def test_cancel_in_flight_praq(f32):
    """A PRAQ request cancelled in flight stops at its next round with
    status 3; its answer is a point of the box with a bracket, and a request
    of the other family beside it converges to the batched answer."""
    from reformulation_amd import praq, workloads
    from reformulation_amd.serve import Server
    w = workloads.praq_er(77, 60000, 400000)
    endless = praq.Parameters.with_auto_step(10**8, 0.0, time_limit=120.0)
    other = [m for m in _members() if m[0] == "pria"][:1]
    b, index = _batch(other, f32)
    b.solve("batched")
    want = _results(b, index)[0]
    b.close()
    with Server(use_float32=f32, host_threads=2) as srv:
        a = srv.submit_praq(w.problem(), endless)
        t = _submit(srv, _problems(other)[0])
        end = time.time() + 60
        while srv.info(a).admitted < 0:
            assert time.time() < end
            time.sleep(0.005)
        time.sleep(0.1)
        assert srv.cancel(a), srv.info(a)
        srv.wait_idle(60)
        x, y, ia = srv.praq_result(a)
        assert ia.status == 3 and ia.iterations > 0
        assert np.all((x >= 0.0) & (x <= 1.0))
        assert ia.lower <= ia.upper + 1e-9 * abs(ia.upper)
        _check_same(_served(srv, "pria", t), want, 0)
