"""The device-pointer entry points (praq_solve_device, pria_solve_device):
instances and answers in caller-owned device memory (torch tensors), work on
the caller's stream.  Marked gpu.

Parity bar.  PRAQ: the answer, iterations and bounds are bitwise those of
praq_solve_ex on the same instance (the same host build and device work; only
where the arrays live differs).  PRIA: the scale, the image and the upper
bound are computed on the device in an order of its own, so the run matches
pria_solve_ex to rounding: bounds within 1e-9 relative, iterations within one
check interval, x within 1e-6; the device upper bound equals numpy's
primal_objective of the returned x to 1e-12 relative, and the lower bound
rebuilt from the returned normalised duals equals info.lower to 1e-9."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(a, dtype=None):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).cuda()


@pytest.mark.parametrize("f32", [False, True])
@pytest.mark.parametrize("size", [(600, 2400), (2000, 8000), (40000, 160000)])
def test_praq_device_entry_equals_the_host_entry(f32, size):
    import torch
    from reformulation_amd import praq, workloads
    n, m = size
    w = workloads.praq_er(31, n, m)
    problem = w.problem()
    params = praq.Parameters.with_auto_step(200_000, 1e-6, use_float32=f32)
    print(f"praq device entry n={n} m={m} f32={f32}", flush=True)
    hx, hy, hinfo = praq.solve_with_info(problem, params)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):   # inputs, outputs and the solve in one stream's order
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        y = torch.zeros((m, 2), dtype=torch.float64, device="cuda")
        info = praq.solve_device(_dev(problem.diag), _dev(problem.rows), _dev(problem.cols),
                                 _dev(problem.vals), params, x, y, stream=stream)
    stream.synchronize()
    assert info.status == hinfo.status == 0
    assert info.iterations == hinfo.iterations and info.tier == hinfo.tier
    assert info.lower == hinfo.lower and info.upper == hinfo.upper
    assert np.array_equal(x.cpu().numpy(), hx)
    assert np.array_equal(y.cpu().numpy(), hy)


def _praq_device_vs_host(problem, params):
    import torch
    from reformulation_amd import praq
    n, m = problem.n(), problem.m()
    hx, hy, hinfo = praq.solve_with_info(problem, params)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        y = torch.zeros((m, 2), dtype=torch.float64, device="cuda") if m else None
        info = praq.solve_device(_dev(problem.diag), _dev(problem.rows, torch.int64),
                                 _dev(problem.cols, torch.int64), _dev(problem.vals), params, x,
                                 y, stream=stream)
    stream.synchronize()
    assert info.status == hinfo.status and info.tier == hinfo.tier
    assert info.iterations == hinfo.iterations and info.scale == hinfo.scale
    assert info.lower == hinfo.lower and info.upper == hinfo.upper
    assert np.array_equal(x.cpu().numpy(), hx)
    if m:
        assert np.array_equal(y.cpu().numpy(), hy)
    return info


@pytest.mark.parametrize("f32", [False, True])
def test_praq_device_build_equals_the_host_build(f32):
    """praq_solve_device builds the half-edge CSR on the device (a stable
    radix sort of the half-edges by node, hub lists, scale, start): a
    power-law graph whose hubs take the chunked gather, the dense-degree mode
    and an edgeless instance solve bit for bit as the host build's."""
    import dataclasses
    from reformulation_amd import praq, workloads
    params = praq.Parameters.with_auto_step(200_000, 1e-6, use_float32=f32)
    cl = workloads.praq_chung_lu(3, 100_000, mean_degree=16.0)
    print(f"device build: chung-lu n={cl.n} m={cl.edges} f32={f32}", flush=True)
    info = _praq_device_vs_host(cl.problem(), params)
    assert info.tier == 2
    er = workloads.praq_er(8, 3000, 12000)
    print("device build: dense degree", flush=True)
    _praq_device_vs_host(er.problem(), dataclasses.replace(params, dense_degree=1))
    edgeless = praq.QuboProblem.from_sparse(5, np.array([1.0, -2.0, 0.5, 0.0, -1.0]),
                                            np.zeros(0, np.int64), np.zeros(0, np.int64),
                                            np.zeros(0))
    print("device build: edgeless", flush=True)
    _praq_device_vs_host(edgeless, params)


def test_praq_device_build_rejects_invalid_edges():
    import torch
    from reformulation_amd import praq
    params = praq.Parameters.with_auto_step(1000, 1e-6)
    diag = _dev(np.zeros(4))
    vals = _dev(np.ones(2))
    x = torch.zeros(4, dtype=torch.float64, device="cuda")
    for rows, cols in (([0, 2], [1, 2]), ([0, 1], [1, 4]), ([-1, 0], [2, 3])):
        print(f"device build: invalid edges {rows} {cols}", flush=True)
        with pytest.raises(ValueError, match="status -2"):
            praq.solve_device(diag, _dev(np.array(rows), torch.int64),
                              _dev(np.array(cols), torch.int64), vals, params, x)


def _pria_case(n, t, seed):
    from reformulation_amd import workloads
    j = workloads.pria_c1(seed, n, t)
    return j.order_costs, j.joint_order_cost, j.holding_costs


@pytest.mark.parametrize("f32", [False, True])
@pytest.mark.parametrize("shape", [(8, 17), (20, 60), (50, 200)])
def test_pria_device_entry_matches_the_host_entry(f32, shape):
    import torch
    from reformulation_amd import pria
    n, t = shape
    k, k0, h = _pria_case(n, t, 41)
    problem = pria.JrpProblem.new(k, k0, h)
    params = pria.Parameters.with_auto_step(200_000, 1e-6, use_float32=f32)
    print(f"pria device entry n={n} t={t} f32={f32}", flush=True)
    hx, hyy, hzz, hinfo = pria.solve_with_info(problem, params)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):   # inputs, outputs and the solve in one stream's order
        x, y, z = (torch.zeros((n, t, t), dtype=torch.float64, device="cuda") for _ in range(3))
        info = pria.solve_device(_dev(k), k0, _dev(h), params, x, y, z, stream=stream)
    stream.synchronize()
    assert info.status == hinfo.status == 0 and info.tier == hinfo.tier
    sc = 1.0 + abs(hinfo.upper) + abs(hinfo.lower)
    assert abs(info.upper - hinfo.upper) <= 1e-9 * sc
    assert abs(info.lower - hinfo.lower) <= 1e-9 * sc
    assert abs(info.iterations - hinfo.iterations) <= params.op_steps
    dx = x.cpu().numpy()
    assert np.max(np.abs(dx - hx)) <= 1e-6
    # the host sums the squares in sequence (rounding up to ~count * eps
    # relative, 2e-10 for 1e6 entries); the device in blocks
    assert abs(info.scale - hinfo.scale) <= 1e-10 * hinfo.scale
    # the device upper bound is the objective of the returned x; the duals
    # come back normalised as the host entry returns them
    assert abs(info.upper - pria.primal_objective(problem, dx)) <= 1e-12 * sc
    lower, _ = pria.relaxation_bounds(problem, dx, y.cpu().numpy(), z.cpu().numpy())
    hlower, _ = pria.relaxation_bounds(problem, hx, hyy, hzz)
    assert abs(lower - hlower) <= 1e-9 * sc
    assert np.all(np.tril(dx[0], -1) == 0.0)


def test_pria_device_entry_without_outputs_still_reports_the_bounds():
    import torch
    from reformulation_amd import pria
    k, k0, h = _pria_case(12, 41, 43)
    params = pria.Parameters.with_auto_step(200_000, 1e-6)
    info = pria.solve_device(_dev(k), k0, _dev(h), params)
    torch.cuda.synchronize()
    x = torch.zeros((12, 41, 41), dtype=torch.float64, device="cuda")
    info2 = pria.solve_device(_dev(k), k0, _dev(h), params, x)
    assert info.status == info2.status == 0
    assert info.iterations == info2.iterations
    assert info.upper == info2.upper and info.lower == info2.lower
