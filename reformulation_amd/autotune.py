"""Autotuner for the kernel shapes (include/reform_tune.h).

    python -m reformulation_amd.autotune [--out reformulation_amd/tuning/mi355x.json]
    python -m reformulation_amd.autotune --batched [--count 1024]

Single-instance device tier (default): for each mean-degree bucket b (mean
degree 2m/n in [2^b, 2^(b+1))) it draws an Erdos-Renyi instance at the
bucket's geometric centre (n nodes, fewer where that keeps it at 4M edges,
`workloads.praq_er` weights), times one reflected device-tier PDHG step with
every lane count 1, 2, ..., 64 (`reform_time_praq_steps`, HIP events, FP32
and FP64 iterates), and keeps the fastest summed over both precisions
("praq.lanes.<b>").

Batched and served tiers (--batched): on the batched mixture itself
(`workloads.mix`, FP32 iterates to 1e-6, every step launch timed by HIP
events), the PRIA members' FP32 x-tile shape (threads 256/512, x1 held in
registers or read again) by the x kernel's summed time, their y/z register
shape (compact/full) by the y/z kernel's, and for each mean-degree bucket
the PRAQ members fall in, the lanes per node by the node kernel's summed
time over that bucket's members ("batch.praq.lanes.<b>", "batch.pria.*").

The table is written as JSON {"table": {key: value, ...}, ...} with the
timings beside it, keeping entries of the other procedure already in the
file; `load()` (called when the package loads the library) puts a table
into the library.
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
    L.reform_batch_praq_lanes_for.restype = ctypes.c_int
    L.reform_batch_praq_lanes_for.argtypes = [ctypes.c_int64, ctypes.c_int64]
    L.reform_time_praq_steps.restype = ctypes.c_int
    L.reform_time_praq_steps.argtypes = [ctypes.POINTER(_native.PraqInstance), ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_double)]
    L._tune_declared = True


def apply(table: dict, L=None) -> None:
    """Put every entry of `table` into the library (ValueError for a key or
    value the library rejects)."""
    L = L or _native.lib()
    _declare(L)
    for key, value in table.items():
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


def batch_lanes_for(n: int, m: int) -> int:
    L = _native.lib()
    _declare(L)
    return int(L.reform_batch_praq_lanes_for(n, m))


def table_in_use(L=None) -> dict:
    """The library's current entries of every key (0 = built-in choice)."""
    L = L or _native.lib()
    _declare(L)
    keys = ([f"praq.lanes.{b}" for b in range(8)] + [f"batch.praq.lanes.{b}" for b in range(8)]
            + list(PRIA_KNOBS))
    return {k: int(L.reform_tune_get(k.encode())) for k in keys}


#: the batched PRIA knobs and their candidate values (include/reform_tune.h)
PRIA_KNOBS = {"batch.pria.x_threads": (256, 512), "batch.pria.x_hold": (1, 2),
              "batch.pria.yz_shape": (1, 2)}


def _bucket(n: int, m: int) -> int:
    degree = 2.0 * m / max(n, 1)
    b = 0
    while b + 1 < 8 and degree >= 2.0 ** (b + 1):
        b += 1
    return b


def _kernel_ms(members, tolerance: float, L, repeats: int = 2) -> dict:
    """Summed milliseconds per step kernel of a timed batched solve of
    `members` (the least of `repeats`, after one untimed solve), with the
    library's current table."""
    from . import praq, pria
    from .batch import Batch
    b = Batch(use_float32=True)
    try:
        for kind, inst in members:
            if kind == "praq":
                b.add_praq(inst.problem(), praq.Parameters.with_auto_step(200_000, tolerance))
            else:
                b.add_pria(inst.problem(), pria.Parameters.with_auto_step(200_000, tolerance))
        b.upload()
        b.solve("batched")
        b.set_timing(True)
        best = {}
        for _ in range(repeats):
            b.solve("batched")
            for k in b.kernel_stats():
                best[k["name"]] = min(best.get(k["name"], float("inf")), k["ms"])
        return best
    finally:
        b.close()


def tune_batched(count: int = 1024, tolerance: float = 1e-6, log=print, L=None) -> dict:
    """The batched tiers' entries, measured on the mixture (module docstring);
    returns {"table": {...}, "timings": {...}}.  The library's batched
    entries are cleared first and left at the chosen values."""
    from . import workloads
    L = L or _native.lib()
    _declare(L)
    for key in list(PRIA_KNOBS) + [f"batch.praq.lanes.{b}" for b in range(8)]:
        L.reform_tune_set(key.encode(), 0)
    members = workloads.mix(count)
    table, timings = {}, {}
    # PRIA: x shape by the x kernel, y/z shape by the y/z kernel
    jrp = [mb for mb in members if mb[0] == "pria"]
    if jrp:
        x_ms = {}
        shapes = [(t, h) for t in PRIA_KNOBS["batch.pria.x_threads"]
                  for h in PRIA_KNOBS["batch.pria.x_hold"]]
        # each shape measured in a forward and a backward sweep (the least
        # counts), so a drift of the clocks over the sweep favours none
        for threads, hold in shapes + shapes[::-1]:
            L.reform_tune_set(b"batch.pria.x_threads", threads)
            L.reform_tune_set(b"batch.pria.x_hold", hold)
            ms = _kernel_ms(jrp, tolerance, L)["wpria_x"]
            x_ms[(threads, hold)] = min(x_ms.get((threads, hold), float("inf")), ms)
            log(f"pria x tiles: {threads} threads, hold {hold}: {ms:.1f} ms")
        threads, hold = min(x_ms, key=x_ms.get)
        table["batch.pria.x_threads"], table["batch.pria.x_hold"] = threads, hold
        L.reform_tune_set(b"batch.pria.x_threads", threads)
        L.reform_tune_set(b"batch.pria.x_hold", hold)
        yz_ms = {}
        for shape in PRIA_KNOBS["batch.pria.yz_shape"]:
            L.reform_tune_set(b"batch.pria.yz_shape", shape)
            yz_ms[shape] = _kernel_ms(jrp, tolerance, L)["wpria_yz"]
            log(f"pria y/z shape {shape}: {yz_ms[shape]:.1f} ms")
        table["batch.pria.yz_shape"] = min(yz_ms, key=yz_ms.get)
        L.reform_tune_set(b"batch.pria.yz_shape", table["batch.pria.yz_shape"])
        timings["batch.pria.x"] = {f"{t}/{h}": v for (t, h), v in x_ms.items()}
        timings["batch.pria.yz_shape"] = {str(k): v for k, v in yz_ms.items()}
    # PRAQ: lanes per bucket, by the node kernel over the bucket's members
    qubo = [mb for mb in members if mb[0] == "praq"]
    buckets = sorted({_bucket(inst.n, inst.edges) for _, inst in qubo})
    for b in buckets:
        group = [mb for mb in qubo if _bucket(mb[1].n, mb[1].edges) == b]
        key = f"batch.praq.lanes.{b}"
        row = {}
        for lanes in LANES:
            if lanes > 2 ** (b + 2):      # > 4 lanes per half-edge at the bucket's top
                break
            L.reform_tune_set(key.encode(), lanes)
            row[lanes] = _kernel_ms(group, tolerance, L)["wpraq_node"]
            log(f"praq bucket {b} ({len(group)} instances): {lanes} lanes {row[lanes]:.1f} ms")
        table[key] = min(row, key=row.get)
        L.reform_tune_set(key.encode(), table[key])
        timings[key] = dict(instances=len(group), ms={str(k): v for k, v in row.items()})
    return {"table": table, "timings": timings}


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
    ap.add_argument("--batched", action="store_true",
                    help="tune the batched tiers on the mixture instead")
    ap.add_argument("--count", type=int, default=1024, help="mixture size (--batched)")
    args = ap.parse_args()
    L = _native.lib()
    _declare(L)
    log = lambda s: print(s, file=sys.stderr, flush=True)   # noqa: E731
    out = Path(args.out)
    old = json.loads(out.read_text()) if out.exists() else {}
    if args.batched:
        result = tune_batched(args.count, log=log, L=L)
        keep = {k: v for k, v in old.get("table", {}).items() if not k.startswith("batch.")}
        merged = dict(old, table={**keep, **result["table"]})
        merged["batched_timings"] = result["timings"]
        merged["batched_rule"] = (f"batched mixture of {args.count} instances, FP32 iterates "
                                  "to 1e-6: PRIA x shape by the x kernel's summed ms, y/z "
                                  "shape by the y/z kernel's, PRAQ lanes per mean-degree "
                                  "bucket by the node kernel's over the bucket's members")
    else:
        # measure against the built-in rule, not an old table
        for b in range(8):
            L.reform_tune_set(f"praq.lanes.{b}".encode(), 0)
        lo, _, hi = args.buckets.partition("-")
        result = tune(args.n, range(int(lo), int(hi or lo) + 1), log=log)
        keep = {k: v for k, v in old.get("table", {}).items() if k.startswith("batch.")}
        merged = dict(old, table={**result["table"], **keep}, timings=result["timings"])
        merged["rule"] = ("lanes per node for PRAQ mean-degree buckets [2^b, 2^(b+1)); one "
                          "reflected device-tier step, FP32 + FP64 ms summed")
    merged["device"] = "MI355X (gfx950)"
    out.parent.mkdir(parents=True, exist_ok=True)
    out.write_text(json.dumps(merged, indent=1) + "\n")
    print(json.dumps(merged["table"]), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
