"""The C-ABI library builds for gfx950, loads without a GPU and exports every
symbol include/*.h declares; the host-side records match.  CPU only."""
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for header in (ROOT / "include").glob("*.h"):
        text = header.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b([a-z_][a-z0-9_]*)\s*\(", text, flags=re.M):
            name = m.group(1)
            if name not in ("if", "while", "for", "return", "sizeof"):
                names.add(name)
    return names


def test_every_declared_symbol_is_exported():
    from reformulation_amd import _native
    lib = _native.lib()
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_native.EXPORTS) <= declared_symbols()


def test_headers_declare_the_reference_abi():
    names = declared_symbols()
    for n in ("praq_default_parameters", "praq_check_interval", "praq_solve", "pria_params_auto",
              "pria_solve_f64", "pria_solve_f32", "pria_solve_batch_f64", "pria_solve_batch_f32",
              "pria_objective_f64"):
        assert n in names


def test_host_only_entry_points_without_a_gpu():
    """Entry points that do no device work run on any host."""
    import ctypes
    import numpy as np
    from reformulation_amd import _native
    lib = _native.lib()
    p = lib.praq_default_parameters()
    assert p.tolerance == 1e-8 and p.op_steps == 16 and p.max_iter == -1
    assert lib.praq_check_interval(400_000, ctypes.byref(p)) == 9
    assert lib.praq_check_interval(2_000_000, ctypes.byref(p)) == 16
    q = lib.pria_params_auto(-1, 1e-6)
    assert abs(q.eta - 0.999 / 2 ** 0.5) < 1e-15 and q.op_steps == 16 and q.omega == 0.10835
    # pria_objective_f64 on the host: all mass on the diagonal s = t
    n, t = 2, 3
    h = np.zeros((n, t, t))
    x = np.zeros((n, t, t))
    for i in range(n):
        for s in range(t):
            x[i, s, s] = 1.0
    k = np.array([2.0, 3.0])
    val = lib.pria_objective_f64(n, t, _native.dptr(k), 5.0, _native.dptr(h), _native.dptr(x))
    assert val == pytest.approx(t * k.sum() + 5.0 * t)


def test_praq_rejects_invalid_arguments_before_any_device_work():
    import ctypes
    import numpy as np
    from reformulation_amd import _native
    lib = _native.lib()
    p = lib.praq_default_parameters()
    res = _native.PraqResult()
    x = np.zeros(3)
    diag = np.zeros(3)
    rows = np.array([0], dtype=np.int64)
    cols = np.array([0], dtype=np.int64)
    vals = np.array([1.0])
    assert lib.praq_solve(None, ctypes.byref(p), _native.dptr(x), ctypes.byref(res)) == -1
    inst = _native.PraqInstance(3, 1, _native.dptr(diag), _native.iptr(rows), _native.iptr(cols),
                                _native.dptr(vals))
    assert lib.praq_solve(ctypes.byref(inst), ctypes.byref(p), None, ctypes.byref(res)) == -1
    # a self-loop, an index past n and n = 0 are instance errors (-2)
    assert lib.praq_solve(ctypes.byref(inst), ctypes.byref(p), _native.dptr(x), ctypes.byref(res)) == -2
    cols[0] = 3
    assert lib.praq_solve(ctypes.byref(inst), ctypes.byref(p), _native.dptr(x), ctypes.byref(res)) == -2
    inst.n = 0
    cols[0] = 1
    assert lib.praq_solve(ctypes.byref(inst), ctypes.byref(p), _native.dptr(x), ctypes.byref(res)) == -2
    bad = lib.praq_default_parameters()
    bad.omega = -1.0
    inst.n = 3
    assert lib.praq_solve(ctypes.byref(inst), ctypes.byref(bad), _native.dptr(x), ctypes.byref(res)) == -1


def test_device_pointer_entries_validate_before_any_device_work():
    """praq_solve_device / pria_solve_device reject bad shapes and parameters
    before they touch a pointer (the pointers here are not device memory)."""
    import ctypes
    from reformulation_amd import _native
    lib = _native.lib()
    p = lib.praq_default_parameters()
    o = lib.praq_default_options()
    info = _native.PraqInfo()
    fake = ctypes.cast(ctypes.c_void_p(4096), _native._dp)
    ifake = ctypes.cast(ctypes.c_void_p(4096), _native._ip)
    inst = _native.PraqInstance(0, 1, fake, ifake, ifake, fake)
    call = lambda inst, p, x: lib.praq_solve_device(ctypes.byref(inst), ctypes.byref(p),
                                                    ctypes.byref(o), x, None,
                                                    ctypes.byref(info), None)
    assert call(inst, p, 4096) == -2                       # n = 0
    inst.n = 3
    assert call(inst, p, None) == -1                       # no x
    inst.rows = None
    assert call(inst, p, 4096) == -2                       # m > 0 without rows
    bad = lib.praq_default_parameters()
    bad.omega = -1.0
    assert call(inst, bad, 4096) == -1
    assert info.status == -1 and info.lower != info.lower   # NaN bounds on error
    q = lib.pria_params_auto(100, 1e-6)
    po = lib.pria_default_options()
    pinfo = _native.PriaInfo()
    pcall = lambda n, t, q: lib.pria_solve_device(n, t, 4096, 1.0, 4096, ctypes.byref(q),
                                                  ctypes.byref(po), 0, 4096, None, None,
                                                  ctypes.byref(pinfo), None)
    assert pcall(0, 10, q) == -1
    assert pcall(2, 1201, q) == -1                         # T' above 1200
    q.omega = 0.0
    assert pcall(2, 10, q) == -1
    assert lib.pria_solve_device(2, 10, None, 1.0, 4096, ctypes.byref(q), None, 0, None, None,
                                 None, None, None) == -1


def test_no_alloc_entry_validates_and_sizes_before_any_device_work():
    """pria_solve_device_noalloc rejects what pria_solve_device rejects, and an
    instance its static array cannot hold, with an explanation, before any
    HIP call (the pointers here are not device memory)."""
    import ctypes
    from reformulation_amd import _native
    lib = _native.lib()
    q = lib.pria_params_auto(100, 1e-6)
    info = _native.PriaInfo()
    call = lambda n, t, q, f32=0: lib.pria_solve_device_noalloc(
        n, t, 4096, 1.0, 4096, ctypes.byref(q), None, f32, 4096, None, None,
        ctypes.byref(info), None)
    assert call(0, 10, q) == -1
    assert call(2, 1201, q) == -1                          # T' above 1200
    # 100 x 300: 4.5e6 entries, about 140 bytes each with FP64 iterates
    assert call(100, 300, q) == -1 and info.status == -1
    err = _native.last_error()
    assert "needs" in err and str(64 << 20) in err
    # 30 x 200 (603000 entries) with FP32 iterates and room for their FP64
    # handover image (about 210 bytes an entry): some 120 MiB
    assert call(30, 200, q, 1) == -1 and "needs" in _native.last_error()
    bad = lib.pria_params_auto(100, 1e-6)
    bad.omega = 0.0
    assert call(2, 10, bad) == -1
    assert lib.pria_solve_device_noalloc(2, 10, None, 1.0, 4096, ctypes.byref(q), None, 0, None,
                                         None, None, None, None) == -1


def test_batch_validates_on_the_host_before_any_device_work():
    import numpy as np
    from reformulation_amd import praq, pria
    from reformulation_amd.batch import Batch
    b = Batch(use_float32=True)
    assert b.counts() == (0, 0)
    ok = praq.QuboProblem.from_sparse(3, np.zeros(3), np.array([0]), np.array([1]),
                                      np.array([1.0]))
    assert b.add_praq(ok, praq.Parameters.with_auto_step(100, 1e-6)) == 0
    k, h = np.array([2.0, 3.0]), np.zeros((2, 4, 4))
    assert b.add_pria(pria.JrpProblem.new(k, 5.0, h), pria.Parameters.with_auto_step(100, 1e-6)) == 0
    assert b.counts() == (1, 1)
    with pytest.raises(RuntimeError):          # nothing solved yet
        b.praq_result(0)
    with pytest.raises(RuntimeError):          # solve before upload
        b.solve("batched")
    # a self-loop is an instance error found by the host build, before any
    # device allocation
    loop = praq.QuboProblem.from_sparse(3, np.zeros(3), np.array([1]), np.array([1]),
                                        np.array([1.0]))
    b.add_praq(loop, praq.Parameters.with_auto_step(100, 1e-6))
    with pytest.raises(RuntimeError, match=r"\(-2\)"):
        b.upload()
    b.set_compaction(False)                    # host-side switches need no device
    b.set_compaction(True)
    assert b._L.reform_batch_set_compaction(None, 1) == -1
    b.close()


def test_file_clis_reject_usage_and_bad_files_before_any_device_work(tmp_path):
    import subprocess
    from reformulation_amd.harness import wire
    for tool in ("praq_solve_file", "pria_solve_file"):
        exe = wire.BIN / tool
        assert exe.exists(), f"{exe} not built"
        assert subprocess.run([str(exe)], capture_output=True).returncode == 2
        bad = tmp_path / f"{tool}.bin"
        bad.write_bytes(b"\x00" * 40)
        r = subprocess.run([str(exe), str(bad), str(tmp_path / "out.bin")], capture_output=True)
        assert r.returncode == 2 and b"is not a" in r.stderr
        assert not (tmp_path / "out.bin").exists()
    r = subprocess.run([str(wire.BIN / "pria_solve_file"), "x", "y", "1e-6", "-1", "bogus"],
                       capture_output=True)
    assert r.returncode == 2 and b"unknown flag" in r.stderr


def test_wire_format_round_trip(tmp_path):
    import numpy as np
    from reformulation_amd.harness import wire
    p = tmp_path / "q.bin"
    wire.write_praq_instance(p, 3, [1.0, -2.0, 0.5], [0, 1], [1, 2], [0.25, -1.0])
    raw = np.fromfile(p, dtype="<i8", count=3)
    assert raw.tolist() == [wire.PRAQ_MAGIC, 3, 2]
    assert p.stat().st_size == 24 + 8 * 3 + 8 * 2 * 3
    q = tmp_path / "j.bin"
    wire.write_pria_instance(q, [2.0, 3.0], 5.0, np.zeros((2, 4, 4)))
    assert q.stat().st_size == 24 + 8 + 16 + 8 * 32


def test_server_validates_arguments_without_a_gpu():
    """The server's entry points reject null handles and bad arguments before
    touching a device; creating one needs a GPU and fails cleanly without."""
    import ctypes
    from reformulation_amd import _native, serve
    lib = _native.lib()
    serve._declare(lib)
    c = lib.reform_server_default_config()
    assert c.use_float32 == 1 and c.hbm_budget == 200 * 10**9 and c.borrow_inputs == 0
    assert c.host_threads > 0 and c.max_in_flight > 0 and c.build_ahead > 0
    p = lib.praq_default_parameters()
    inst = _native.PraqInstance(3, 0, None, None, None, None)
    assert lib.reform_server_submit_praq(None, ctypes.byref(inst), ctypes.byref(p), None) == -1
    assert lib.reform_server_poll(None, None, 0) == -1
    assert lib.reform_server_wait_idle(None, 1.0) == -1
    assert lib.reform_server_info(None, 0, None) == -1
    assert lib.reform_server_release(None, 0) == -1
    assert lib.reform_server_stats(None, None) == -1
    import torch
    if not torch.cuda.is_available():
        with pytest.raises(RuntimeError):
            serve.Server()


def test_fifo_queue_and_poisson_arrivals():
    """The plain solver's queue model and the arrival process, against closed
    forms: arrivals every 1 s with 2 s of service wait q s more each; a
    Poisson process's count over a window has mean and variance rate*window."""
    import numpy as np
    from reformulation_amd import serve
    lat = serve.fifo_latencies(np.arange(5.0), np.full(5, 2.0))
    assert np.array_equal(lat, 2.0 + np.arange(5.0))
    lat = serve.fifo_latencies(np.arange(5.0) * 3.0, np.full(5, 2.0))   # never queued
    assert np.array_equal(lat, np.full(5, 2.0))
    counts = [len(serve.poisson_arrivals(200.0, 1.0, seed)) for seed in range(200)]
    assert abs(np.mean(counts) - 200.0) < 5.0 and abs(np.var(counts) / 200.0 - 1.0) < 0.3
    a = serve.poisson_arrivals(50.0, 2.0, 7)
    assert np.all(np.diff(a) > 0) and a[0] >= 0.0 and a[-1] < 2.0
    s = serve.summarize(np.array([0.0, 1.0, 2.0]), np.array([0.5, 1.5, 4.0]))
    assert s["requests"] == 3 and s["p50_ms"] == pytest.approx(500.0)
    assert s["solves_per_s"] == pytest.approx(3 / 4.0)


def test_tuning_table_round_trip_and_lane_rule_without_a_gpu(tmp_path):
    """The tuning table accepts only known keys and power-of-two lanes, the
    lane rule reads it per mean-degree bucket, and a table file loads."""
    import json
    from reformulation_amd import _native, autotune
    L = _native.lib()
    autotune._declare(L)
    L.reform_tune_clear()
    # built-in rule: smallest power of two with <= 8 half-edges per lane
    assert autotune.lanes_for(1000, 4000) == 1      # mean degree 8
    assert autotune.lanes_for(1000, 16000) == 4     # 32
    assert autotune.lanes_for(1000, 400000) == 64   # 800
    assert L.reform_tune_set(b"praq.lanes.5", 16) == 0
    assert L.reform_tune_get(b"praq.lanes.5") == 16
    assert autotune.lanes_for(1000, 16000) == 16    # mean degree 32: bucket 5
    assert autotune.lanes_for(1000, 15000) == 4     # mean 30: bucket 4, unset
    for bad in ((b"praq.lanes.8", 4), (b"praq.lanes.5", 3), (b"praq.lanes.5", 128),
                (b"pria.tile", 32), (b"praq.lanes.", 4)):
        assert L.reform_tune_set(*bad) == -1
    assert L.reform_tune_get(b"nope") == -1
    table = {"table": {"praq.lanes.0": 2, "praq.lanes.7": 64}}
    path = tmp_path / "t.json"
    path.write_text(json.dumps(table))
    L.reform_tune_clear()
    assert autotune.load(path) == table["table"]
    assert autotune.lanes_for(1000, 100) == 2 and autotune.lanes_for(10, 2000) == 64
    L.reform_tune_clear()
    autotune.load()                                  # the package's own table, if any


def test_batched_tuning_keys_without_a_gpu():
    """The batched tiers' keys: lanes per bucket read by the batch rule
    (the built-in one while unset, the device-tier entries not consulted),
    the PRIA knobs accept exactly their documented values."""
    from reformulation_amd import _native, autotune
    L = _native.lib()
    autotune._declare(L)
    saved = autotune.table_in_use(L)
    try:
        L.reform_tune_clear()
        assert autotune.batch_lanes_for(1000, 16000) == 4          # rule, mean degree 32
        assert L.reform_tune_set(b"praq.lanes.5", 64) == 0         # device tier only
        assert autotune.batch_lanes_for(1000, 16000) == 4
        assert L.reform_tune_set(b"batch.praq.lanes.5", 8) == 0
        assert autotune.batch_lanes_for(1000, 16000) == 8
        assert autotune.batch_lanes_for(1000, 15000) == 4          # bucket 4, unset
        for key, values in autotune.PRIA_KNOBS.items():
            for v in values + (0,):
                assert L.reform_tune_set(key.encode(), v) == 0
                assert L.reform_tune_get(key.encode()) == v
        for bad in ((b"batch.pria.x_threads", 128), (b"batch.pria.x_hold", 3),
                    (b"batch.pria.yz_shape", -1), (b"batch.praq.lanes.9", 4),
                    (b"batch.praq.lanes.3", 6), (b"batch.pria", 1)):
            assert L.reform_tune_set(*bad) == -1
        used = autotune.table_in_use(L)
        assert used["batch.praq.lanes.5"] == 8 and used["praq.lanes.5"] == 64
        assert used["batch.pria.x_threads"] == 0
    finally:
        L.reform_tune_clear()
        autotune.apply({k: v for k, v in saved.items() if v}, L)


def test_ctypes_records_match_the_c_layouts(tmp_path):
    """Every record the Python side hands to the library has the size and
    field offsets of its C struct (compiled from include/*.h with gcc): a
    mismatch would let the library write past a ctypes buffer."""
    import ctypes
    import subprocess
    from reformulation_amd import _native, batch, serve
    pairs = [("PraqInstance", _native.PraqInstance), ("PraqParameters", _native.PraqParameters),
             ("PraqResult", _native.PraqResult), ("PraqOptions", _native.PraqOptions),
             ("PraqInfo", _native.PraqInfo), ("pria_params", _native.PriaParams),
             ("pria_result_info", _native.PriaResultInfo), ("pria_options", _native.PriaOptions),
             ("pria_info", _native.PriaInfo), ("reform_kernel_stats", batch.KernelStats),
             ("reform_server_config", serve.ServerConfig),
             ("reform_request_info", serve.RequestInfo),
             ("reform_server_counters", serve.ServerCounters)]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "praq_hip.h"',
             '#include "pria_hip.h"', '#include "reform_batch.h"', '#include "reform_serve.h"',
             'int main(void) {']
    for cname, cls in pairs:
        lines.append(f'    printf("{cname} %zu\\n", sizeof({cname}));')
        for field, _ in cls._fields_:
            lines.append(f'    printf("{cname}.{field} %zu\\n", offsetof({cname}, {field}));')
    lines += ['    return 0;', '}']
    src = tmp_path / "layouts.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layouts"
    subprocess.run(["gcc", "-std=c11", "-I", str(ROOT / "include"), str(src), "-o", str(exe)],
                   check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True,
                                                       text=True, check=True).stdout.splitlines())
    for cname, cls in pairs:
        assert int(got[cname]) == ctypes.sizeof(cls), cname
        for field, _ in cls._fields_:
            assert int(got[f"{cname}.{field}"]) == getattr(cls, field).offset, (cname, field)


def test_fault_injector_arms_and_disarms_without_a_gpu():
    """include/reform_debug.h: the injector is process state only; arming
    and disarming it touch no device, and the allocation count only grows."""
    import ctypes
    from reformulation_amd import _native
    L = _native.lib()
    L.reform_debug_fail_device_alloc.argtypes = [ctypes.c_int64]
    L.reform_debug_fail_device_alloc.restype = ctypes.c_int
    L.reform_debug_device_allocs.restype = ctypes.c_int64
    before = L.reform_debug_device_allocs()
    assert L.reform_debug_fail_device_alloc(3) == 0
    assert L.reform_debug_fail_device_alloc(-7) == 0
    assert L.reform_debug_device_allocs() == before >= 0
