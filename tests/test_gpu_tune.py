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
