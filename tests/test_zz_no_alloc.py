"""pria_solve_device_noalloc: pria_solve_device with every scratch byte in the
library's static per-GPU array.  Marked gpu; called through the C ABI.

Parity bar: the same kernels on the same inputs, so the answers (x, y, z),
bounds, iteration counts, tier, handover iteration and device byte counts
are bitwise those of pria_solve_device.  Two calls in a row on one GPU, the
second on a smaller instance, so the second runs in an array that still
holds the first call's larger image; then an instance too large for the
array is rejected with an explanation and the array still serves."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_FIELDS = ("iterations", "restarts", "checks", "lower", "upper", "upper_iterate",
           "relative_gap", "omega", "scale", "status", "tier", "switch_iteration",
           "device_bytes")


def _case(n, t, seed):
    from reformulation_amd import workloads
    j = workloads.pria_c1(seed, n, t)
    return j.order_costs, j.joint_order_cost, j.holding_costs


def _solve(name, n, t, k, k0, h, params, want_x):
    """One call of `name` on fresh device copies of the instance; the answers
    and the info record on the host."""
    import torch
    from reformulation_amd import _native
    L = _native.lib()
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a)).cuda()
    kd, hd = dev(k), dev(h)
    outs = [torch.full((n, t, t), -7.0, dtype=torch.float64, device="cuda") for _ in range(3)]
    if not want_x:
        outs[0] = None
    p = params.native()
    o = L.pria_default_options()
    o.tier = params.tier
    o.time_limit = params.time_limit
    info = _native.PriaInfo()
    stream = torch.cuda.current_stream()
    torch.cuda.synchronize()
    ptr = [None if a is None else a.data_ptr() for a in outs]
    status = getattr(L, name)(n, t, kd.data_ptr(), float(k0), hd.data_ptr(), ctypes.byref(p),
                              ctypes.byref(o), int(params.use_float32), ptr[0], ptr[1], ptr[2],
                              ctypes.byref(info), stream.cuda_stream)
    torch.cuda.synchronize()
    host = [None if a is None else a.cpu().numpy() for a in outs]
    return status, host, {f: getattr(info, f) for f in _FIELDS}


def _same(a, b):
    assert a[0] == b[0]
    for u, v in zip(a[1], b[1]):
        assert (u is None) == (v is None)
        if u is not None:
            assert np.array_equal(u, v)
    for f in _FIELDS:
        x, y = a[2][f], b[2][f]
        assert x == y or (x != x and y != y), (f, x, y)


def test_no_alloc_entry_equals_the_device_entry_on_two_calls_in_a_row():
    from reformulation_amd import _native, pria
    # 1: device tier (N T'(T'+1)/2 = 101000 entries), FP32 iterates to a
    # tolerance they cannot reach, so the FP64 handover image is used too
    n, t = 20, 100
    k, k0, h = _case(n, t, 71)
    params = pria.Parameters.with_auto_step(200_000, 1e-8, use_float32=True)
    print(f"no-alloc call 1: n={n} t={t} f32 tol=1e-8", flush=True)
    ref = _solve("pria_solve_device", n, t, k, k0, h, params, True)
    got = _solve("pria_solve_device_noalloc", n, t, k, k0, h, params, True)
    assert got[0] in (0, 1)
    assert got[2]["tier"] == 2 and got[2]["switch_iteration"] >= 0
    _same(got, ref)
    # 2: smaller, odd t (padded image), one workgroup, FP64, no x taken
    n, t = 6, 31
    k, k0, h = _case(n, t, 72)
    params = pria.Parameters.with_auto_step(200_000, 1e-6)
    print(f"no-alloc call 2: n={n} t={t} f64 no x", flush=True)
    ref = _solve("pria_solve_device", n, t, k, k0, h, params, False)
    got = _solve("pria_solve_device_noalloc", n, t, k, k0, h, params, False)
    assert got[0] == 0 and got[2]["tier"] == 1
    _same(got, ref)
    # the upper bound still belongs to the answer pria_solve_ex returns
    problem = pria.JrpProblem.new(k, k0, h)
    hx, _, _, hinfo = pria.solve_with_info(problem, params)
    sc = 1.0 + abs(hinfo.upper) + abs(hinfo.lower)
    assert abs(got[2]["upper"] - hinfo.upper) <= 1e-9 * sc


def test_no_alloc_entry_rejects_what_does_not_fit_and_serves_on():
    import torch
    from reformulation_amd import _native, pria
    L = _native.lib()
    # 100 x 300: 4.5e6 entries, far beyond the array
    n, t = 100, 300
    p = pria.Parameters.with_auto_step(1000, 1e-6).native()
    info = _native.PriaInfo()
    h = torch.zeros((n, t, t), dtype=torch.float64, device="cuda")
    k = torch.ones(n, dtype=torch.float64, device="cuda")
    print("no-alloc oversize call", flush=True)
    status = L.pria_solve_device_noalloc(n, t, k.data_ptr(), 1.0, h.data_ptr(), ctypes.byref(p),
                                         None, 0, None, None, None, ctypes.byref(info), None)
    assert status == -1 and info.status == -1
    assert "needs" in _native.last_error()
    del h
    n, t = 8, 17
    kk, k0, hh = _case(n, t, 73)
    params = pria.Parameters.with_auto_step(200_000, 1e-6, use_float32=True)
    print(f"no-alloc call after the rejection: n={n} t={t}", flush=True)
    ref = _solve("pria_solve_device", n, t, kk, k0, hh, params, True)
    got = _solve("pria_solve_device_noalloc", n, t, kk, k0, hh, params, True)
    assert got[0] == 0
    _same(got, ref)
