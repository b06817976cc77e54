"""The file CLIs on the GPU: an instance file in, a result file and
`key: value` stdout out, as the reference harness backends expect
(qubo/harness/backends/praq_cuda.py, jrp/harness/backends/pria_cuda.py).
Marked gpu."""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
from helpers import small_jrp

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_praq_cli_config_instance_matches_highs(tmp_path):
    from reformulation_amd import workloads
    from reformulation_amd.harness import wire
    row = json.loads((GOLDEN / "praq_lp.json").read_text())[0]
    inst = workloads.praq_er(row["seed"], row["n"], row["m"])
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    wire.write_praq_instance(src, inst.n, inst.diag, inst.rows, inst.cols, inst.vals)
    for extra in ([], ["f32"]):
        r = subprocess.run([str(wire.BIN / "praq_solve_file"), str(src), str(dst), "1e-6", "0",
                            *extra], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        res = wire.read_praq_result(dst)
        kv = wire.parse_stdout(r.stdout)
        assert res["status"] == 0 and kv["status"] == "converged"
        assert int(kv["iterations"]) == res["iterations"] and res["n"] == inst.n
        assert kv["precision"] == ("float32" if extra else "float64")
        opt = row["highs"]
        assert res["lower"] <= opt + 1e-6 * abs(opt) and opt <= res["upper"] + 1e-6 * abs(opt)
        assert float(kv["upper"]) == res["upper"]
        assert res["x"].shape == (inst.n,) and np.all((res["x"] >= 0) & (res["x"] <= 1))


def test_pria_cli_matches_the_oracle_with_solution(tmp_path):
    from reformulation_amd.harness import wire
    k, k0, h = small_jrp(11, 6, 13)
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    wire.write_pria_instance(src, k, k0, h)
    r = subprocess.run([str(wire.BIN / "pria_solve_file"), str(src), str(dst), "1e-8", "-1",
                        "solution"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    res = wire.read_pria_result(dst, with_solution=True)
    o = oracle.pria_solve(k, k0, h, oracle.pria_params(tolerance=1e-8))
    assert res["iterations"] == o.iterations and res["status"] == 0
    sc = 1.0 + abs(o.upper)
    # upper is pria_solve_f64's primal_objective: the padded problem's
    # objective at the returned point, as the reference CLI writes it
    assert abs(res["upper"] - o.upper_iterate) <= 1e-9 * sc
    assert abs(res["lower"] - o.lower) <= 1e-9 * sc
    assert np.max(np.abs(res["x"] - o.x)) <= 1e-8
    kv = wire.parse_stdout(r.stdout)
    assert kv["status"] == "converged" and int(kv["items"]) == 6 and int(kv["periods"]) == 13
    # an iteration cap: status 1, exit 1, the bracket still written
    r = subprocess.run([str(wire.BIN / "pria_solve_file"), str(src), str(dst), "1e-12", "32"],
                       capture_output=True, text=True, timeout=120)
    res = wire.read_pria_result(dst)
    assert r.returncode == 1 and res["status"] == 1 and res["iterations"] == 32
    assert res["lower"] <= o.upper + 1e-9 * sc <= res["upper"] + 2e-9 * sc
