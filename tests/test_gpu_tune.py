"""The autotuner's timing entry point and the tuning table on the GPU.
Marked gpu.

A tuned lane count changes only the order of the CSR gather's sums: the
device tier with any table entry keeps the oracle's iteration counts and
bounds to 1e-9 (FP64)."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def test_step_timing_for_every_lane_count():
    from reformulation_amd import autotune, workloads
    inst = workloads.praq_er(3, 20000, 80000)
    for lanes in autotune.LANES:
        for f32 in (True, False):
            ms = autotune.time_steps(inst, lanes, f32, warmup=1, steps=3)
            assert 0.0 < ms < 100.0, (lanes, f32, ms)


@pytest.mark.parametrize("lanes", [1, 8, 64])
def test_tuned_lanes_keep_the_oracle_iterations(lanes):
    from reformulation_amd import _native, autotune, praq, workloads
    L = _native.lib()
    autotune._declare(L)
    inst = workloads.praq_er(11, 3000, 24000)    # mean degree 16: bucket 4
    o = oracle.praq_solve(inst.n, inst.diag, inst.rows, inst.cols, inst.vals,
                          oracle.praq_params(tolerance=1e-7))
    saved = L.reform_tune_get(b"praq.lanes.4")
    try:
        assert L.reform_tune_set(b"praq.lanes.4", lanes) == 0
        assert autotune.lanes_for(inst.n, inst.edges) == lanes
        params = praq.Parameters.with_auto_step(-1, 1e-7, tier=praq.TIER_DEVICE)
        x, _, info = praq.solve_with_info(inst.problem(), params)
    finally:
        L.reform_tune_set(b"praq.lanes.4", saved)
    assert info.tier == praq.TIER_DEVICE and info.status == 0
    assert info.iterations == o.iterations and info.restarts == o.restarts
    sc = 1.0 + abs(o.upper)
    assert abs(info.upper - o.upper) <= 1e-9 * sc and abs(info.lower - o.lower) <= 1e-9 * sc
    assert np.max(np.abs(x - o.x)) <= 1e-7


# The batched tiers' table entries (include/reform_tune.h), each against the
# fixtures at workload sizes: the reference solver's PRIA run at N x T =
# 50 x 200 (its z slices take the y/z kernel's memory path) and the oracle's
# PRAQ run at n = 2e4, mean degree 15 (bucket 3).
BATCH_ENTRIES = [("batch.pria.x_threads", 512), ("batch.pria.x_hold", 2),
                 ("batch.pria.yz_shape", 2), ("batch.praq.lanes.3", 1),
                 ("batch.praq.lanes.3", 16)]


def _batched(L, entry, use_float32):
    from reformulation_amd import praq, pria, workloads
    from reformulation_amd.batch import Batch
    import json
    from pathlib import Path
    golden = Path(__file__).parent / "golden"
    prow = json.loads((golden / "pria_ref_large.json").read_text())[0]
    qrow = json.loads((golden / "praq_c3.json").read_text())[0]
    jrp = workloads.pria_c1(prow["seed"], prow["items"], prow["periods"])
    qubo = workloads.praq_er(qrow["seed"], qrow["n"], qrow["m"])
    key, value = entry if entry else (None, 0)
    saved = L.reform_tune_get(key.encode()) if key else 0
    try:
        if key:
            assert L.reform_tune_set(key.encode(), value) == 0
        b = Batch(use_float32=use_float32)
        r = b.add_pria(jrp.problem(), pria.Parameters.with_auto_step(None, 1e-6))
        q = b.add_praq(qubo.problem(), praq.Parameters.with_auto_step(None, 1e-6))
        b.upload()
        b.solve("batched")
        out = (b.pria_result(r), b.praq_result(q))
        b.close()
    finally:
        if key:
            L.reform_tune_set(key.encode(), saved)
    return out, prow, qrow


@pytest.mark.parametrize("entry", BATCH_ENTRIES, ids=[f"{k}={v}" for k, v in BATCH_ENTRIES])
def test_batched_table_entries_keep_the_reference_iterations(entry):
    from reformulation_amd import _native, autotune
    L = _native.lib()
    autotune._declare(L)
    print(f"batched FP64 with {entry}", flush=True)
    ((x, y, z, rinfo), (qx, _, qinfo)), prow, qrow = _batched(L, entry, False)
    assert rinfo.status == 0 and rinfo.iterations == prow["iterations"]
    sc = 1.0 + abs(prow["primal"])
    assert abs(rinfo.upper_iterate - prow["primal"]) <= 1e-9 * sc
    assert abs(rinfo.lower - prow["dual"]) <= 1e-9 * sc
    assert qinfo.status == 0
    assert qinfo.iterations == qrow["iterations"] and qinfo.restarts == qrow["restarts"]
    sc = 1.0 + abs(qrow["upper"])
    assert abs(qinfo.upper - qrow["upper"]) <= 1e-9 * sc
    assert abs(qinfo.lower - qrow["lower"]) <= 1e-9 * sc
    # FP32 iterates: converged; the register/memory-placement entries give
    # the unset table's bits
    print(f"batched FP32 with {entry}", flush=True)
    ((x32, _, _, r32), (qx32, _, q32)), _, _ = _batched(L, entry, True)
    assert r32.status == 0 and q32.status == 0
    if entry[0] in ("batch.pria.x_hold", "batch.pria.yz_shape"):
        ((x0, _, _, r0), _), _, _ = _batched(L, None, True)
        assert r32.iterations == r0.iterations and r32.upper == r0.upper
        assert np.array_equal(x32, x0)


def test_sequential_device_tier_keeps_its_hub_lists_under_a_batched_lanes_entry():
    """A batch's sequential (plain) mode runs the device tier with the lanes
    the instance's hub lists were built for, whatever the batched tiers'
    entry: a power-law graph with hubs of degree above 64 lanes keeps the
    oracle's iterations and bounds."""
    from reformulation_amd import _native, autotune, praq, workloads
    from reformulation_amd.batch import Batch
    L = _native.lib()
    autotune._declare(L)
    inst = workloads.praq_chung_lu(5, n=20000, mean_degree=10.0)
    degree = np.bincount(np.concatenate([inst.rows, inst.cols]), minlength=inst.n)
    assert degree.max() > 64 * 16 and (degree > 64).sum() > 10
    o = oracle.praq_solve(inst.n, inst.diag, inst.rows, inst.cols, inst.vals,
                          oracle.praq_params(tolerance=1e-7))
    key = b"batch.praq.lanes.3"                     # mean degree 10
    saved = L.reform_tune_get(key)
    try:
        assert L.reform_tune_set(key, 16) == 0
        b = Batch(use_float32=False)
        q = b.add_praq(inst.problem(),
                       praq.Parameters.with_auto_step(-1, 1e-7, tier=praq.TIER_DEVICE))
        b.upload()
        b.solve("sequential")
        _, _, info = b.praq_result(q)
        b.close()
    finally:
        L.reform_tune_set(key, saved)
    assert info.status == 0
    assert info.iterations == o.iterations and info.restarts == o.restarts
    sc = 1.0 + abs(o.upper)
    assert abs(info.upper - o.upper) <= 1e-9 * sc and abs(info.lower - o.lower) <= 1e-9 * sc
