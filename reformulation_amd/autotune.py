"""Autotuner for the kernel shapes (include/reform_tune.h).

    python -m reformulation_amd.autotune [--out reformulation_amd/tuning/mi355x.json]

For each mean-degree bucket b (mean degree 2m/n in [2^b, 2^(b+1))) it draws
an Erdos-Renyi instance at the bucket's geometric centre (n nodes, fewer
where that keeps it at 4M edges, `workloads.praq_er` weights), times one reflected device-tier PDHG step with
every lane count 1, 2, ..., 64 (`reform_time_praq_steps`, HIP events, FP32
and FP64 iterates), and keeps the fastest summed over both precisions.  The
table is written as JSON {"praq.lanes.<b>": lanes, ...} with the timings
beside it; `load()` (called when the package loads the library) puts a
table into the library.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
from pathlib import Path

from . import _native

DEFAULT_TABLE = Path(__file__).resolve().parent / "tuning" / "mi355x.json"
LANES = (1, 2, 4, 8, 16, 32, 64)


def _declare(L):
    if getattr(L, "_tune_declared", False):
        return
    L.reform_tune_set.restype = ctypes.c_int
    L.reform_tune_set.argtypes = [ctypes.c_char_p, ctypes.c_int64]
    L.reform_tune_get.restype = ctypes.c_int64
    L.reform_tune_get.argtypes = [ctypes.c_char_p]
    L.reform_tune_clear.restype = None
    L.reform_tune_clear.argtypes = []
    L.reform_praq_lanes_for.restype = ctypes.c_int
    L.reform_praq_lanes_for.argtypes = [ctypes.c_int64, ctypes.c_int64]
    L.reform_time_praq_steps.restype = ctypes.c_int
    L.reform_time_praq_steps.argtypes = [ctypes.POINTER(_native.PraqInstance), ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_double)]
    L._tune_declared = True


def apply(table: dict, L=None) -> None:
    """Put the "praq.lanes.<b>" entries of `table` into the library."""
    L = L or _native.lib()
    _declare(L)
    for key, value in table.items():
        if not key.startswith("praq.lanes."):
            continue
        if L.reform_tune_set(key.encode(), int(value)) != 0:
            raise ValueError(f"bad tuning entry {key} = {value}")


def load(path: str | os.PathLike | None = None, L=None) -> dict:
    """Read a table written by `tune` and apply it; {} if the file is absent.
    The default table is REFORM_TUNING (a path, or "off") or the one in the
    package."""
    if path is None:
        env = os.environ.get("REFORM_TUNING")
        if env == "off":
            return {}
        path = env or DEFAULT_TABLE
    p = Path(path)
    if not p.exists():
        return {}
    table = json.loads(p.read_text()).get("table", {})
    apply(table, L)
    return table


def lanes_for(n: int, m: int) -> int:
    L = _native.lib()
    _declare(L)
    return int(L.reform_praq_lanes_for(n, m))


def time_steps(instance, lanes: int, use_float32: bool, warmup: int = 5, steps: int = 20) -> float:
    """Milliseconds per reflected device-tier step of a QuboInstance."""
    L = _native.lib()
    _declare(L)
    inst = _native.PraqInstance(instance.n, instance.edges, _native.dptr(instance.diag),
                                _native.iptr(instance.rows), _native.iptr(instance.cols),
                                _native.dptr(instance.vals))
    ms = ctypes.c_double()
    st = L.reform_time_praq_steps(ctypes.byref(inst), int(use_float32), lanes, warmup, steps,
                                  ctypes.byref(ms))
    if st != 0:
        raise RuntimeError(f"reform_time_praq_steps failed ({st}): {_native.last_error()}")
    return ms.value


def tune(n: int = 200_000, buckets=range(8), seed: int = 7, log=print) -> dict:
    """Measure every lane count on each bucket's instance; returns
    {"table": {...}, "timings": {...}}."""
    from . import workloads
    table, timings = {}, {}
    for b in buckets:
        degree = 2.0 ** (b + 0.5)
        nb = min(n, int(8e6 / degree))       # at most 4M edges
        m = int(round(nb * degree / 2))
        inst = workloads.praq_er(seed + b, nb, m)
        row = {}
        for lanes in LANES:
            row[lanes] = {prec: time_steps(inst, lanes, prec == "fp32")
                          for prec in ("fp32", "fp64")}
        best = min(LANES, key=lambda k: row[k]["fp32"] + row[k]["fp64"])
        table[f"praq.lanes.{b}"] = best
        timings[f"praq.lanes.{b}"] = dict(mean_degree=degree, n=nb, m=m, best=best,
                                          ms={str(k): v for k, v in row.items()})
        log(f"bucket {b} (mean degree {degree:.1f}, m={m}): best lanes {best}; "
            + ", ".join(f"{k}: {v['fp32']:.4f}/{v['fp64']:.4f} ms" for k, v in row.items()))
    return {"table": table, "timings": timings}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(DEFAULT_TABLE))
    ap.add_argument("--n", type=int, default=200_000)
    ap.add_argument("--buckets", default="0-7")
    args = ap.parse_args()
    lo, _, hi = args.buckets.partition("-")
    buckets = range(int(lo), int(hi or lo) + 1)
    L = _native.lib()
    _declare(L)
    L.reform_tune_clear()          # measure against the built-in rule, not an old table
    result = tune(args.n, buckets, log=lambda s: print(s, file=sys.stderr, flush=True))
    result["device"] = "MI355X (gfx950)"
    result["rule"] = ("lanes per node for PRAQ mean-degree buckets [2^b, 2^(b+1)); one "
                      "reflected device-tier step, FP32 + FP64 ms summed")
    Path(args.out).parent.mkdir(parents=True, exist_ok=True)
    Path(args.out).write_text(json.dumps(result, indent=1) + "\n")
    print(json.dumps(result["table"]), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
