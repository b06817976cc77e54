"""Device allocations refused part way through each entry point, through the
fault injector of include/reform_debug.h.  Marked gpu.

For every allocation k a call makes, the call is repeated with the k-th one
refused (as hipMalloc refuses when the device is out of memory): it either
fails with its error code (raised by the Python layer) or, where the refused
memory was optional, completes; any answer it returns, and the answer of the
next call without a fault, are bitwise those of a call that never saw one.
Every call is a fresh handle, so a failed one is also destroyed while its
failure is fresh: the batch and the server drain their streams first."""
import numpy as np
import pytest

from test_gpu_batch import _same_bits

pytestmark = pytest.mark.gpu


def _lib():
    from reformulation_amd import _native
    L = _native.lib()
    L.reform_debug_fail_device_alloc.argtypes = [_native.ctypes.c_int64]
    L.reform_debug_fail_device_alloc.restype = _native.ctypes.c_int
    L.reform_debug_device_allocs.argtypes = []
    L.reform_debug_device_allocs.restype = _native.ctypes.c_int64
    return L


def _sweep(name, call, extra=0):
    """Returns how many of the refused allocations made the call fail."""
    L = _lib()
    ref = call()
    before = L.reform_debug_device_allocs()
    again = call()
    per_call = L.reform_debug_device_allocs() - before
    assert _same_bits(again, ref), name
    assert per_call >= 1, name
    failed = 0
    for k in range(per_call + extra):
        print(f"{name}: allocation {k} of {per_call} refused", flush=True)
        L.reform_debug_fail_device_alloc(k)
        try:
            out = call()
        except (RuntimeError, ValueError):
            failed += 1
        else:
            assert _same_bits(out, ref), (name, k)
        finally:
            L.reform_debug_fail_device_alloc(-1)
    assert _same_bits(call(), ref), name
    return failed


def _flat(x, y, info):
    return (x, y, np.array([info.iterations, info.lower, info.upper]))


@pytest.mark.parametrize("tier", [1, 2])
def test_plain_praq_solve_fails_cleanly(tier):
    from reformulation_amd import praq, workloads
    w = workloads.praq_er(41, 3000, 9000)
    p = w.problem()
    params = praq.Parameters.with_auto_step(50_000, 1e-6, tier=tier)
    assert _sweep(f"praq tier {tier}", lambda: _flat(*praq.solve_with_info(p, params))) >= 1


@pytest.mark.parametrize("tier", [1, 2])
@pytest.mark.parametrize("f32", [False, True])
def test_plain_pria_solve_fails_cleanly(tier, f32):
    from reformulation_amd import pria, workloads
    j = workloads.pria_c1(42, 12, 41)
    prob = pria.JrpProblem.new(j.order_costs, j.joint_order_cost, j.holding_costs)
    params = pria.Parameters.with_auto_step(50_000, 1e-6, use_float32=f32, tier=tier)

    def call():
        x, y, z, info = pria.solve_with_info(prob, params)
        return (x, y, z, np.array([info.iterations, info.lower, info.upper]))
    assert _sweep(f"pria tier {tier} f32 {f32}", call) >= 1


def test_praq_device_entry_fails_cleanly():
    import torch
    from reformulation_amd import praq, workloads
    w = workloads.praq_er(43, 2000, 8000)
    prob = w.problem()
    params = praq.Parameters.with_auto_step(50_000, 1e-6)
    dev = [torch.as_tensor(np.ascontiguousarray(a)).cuda()
           for a in (prob.diag, prob.rows, prob.cols, prob.vals)]

    def call():
        x = torch.zeros(prob.n(), dtype=torch.float64, device="cuda")
        y = torch.zeros((prob.m(), 2), dtype=torch.float64, device="cuda")
        info = praq.solve_device(*dev, params, x, y)
        torch.cuda.synchronize()
        return (x.cpu().numpy(), y.cpu().numpy(), np.array([info.iterations, info.upper]))
    assert _sweep("praq device entry", call) >= 1


@pytest.mark.parametrize("f32", [False, True])
def test_batch_fails_cleanly(f32):
    from test_gpu_batch import _members, _batch_open, _results
    members = _members()[:2] + [m for m in _members() if m[0] == "pria"][:2]

    def call():
        b, index = _batch_open(members, f32)
        try:
            b.upload()
            b.solve("batched")
            out = _results(b, index)
        finally:
            b.close()
        return tuple((x, d, np.array([i.iterations, i.lower, i.upper])) for x, d, i in out)
    assert _sweep(f"batch f32 {f32}", call) >= 1


def test_server_fails_cleanly():
    from reformulation_amd.serve import Server
    from test_gpu_batch import _members
    from test_gpu_serve import _problems, _submit, _served
    items = _problems(_members()[:1] + [m for m in _members() if m[0] == "pria"][:1])

    def call():
        srv = Server(use_float32=True, hbm_budget=4 * 10**9, max_in_flight=64)
        try:
            tickets = [_submit(srv, it) for it in items]
            srv.wait_idle(120)
            out = []
            for (kind, _, _), t in zip(items, tickets):
                x, d, i = _served(srv, kind, t)
                out.append((x, d, np.array([i.status, i.iterations, i.lower, i.upper])))
        finally:
            srv.close()
        return tuple(out)
    # admissions may come in one table build or two: a refused allocation
    # past the first call's count is tried too
    assert _sweep("server", call, extra=2) >= 1
