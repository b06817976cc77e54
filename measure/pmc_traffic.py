"""HBM traffic per step launch of the batched step kernels, from rocprofv3
counter passes, calibrated on a device copy of known size.

    rocprofv3 --pmc FETCH_SIZE --output-format csv -d D/copy_fetch -o c -- python3 -m measure.pmc_copy
    rocprofv3 --pmc WRITE_SIZE --output-format csv -d D/copy_write -o c -- python3 -m measure.pmc_copy
    rocprofv3 --pmc FETCH_SIZE --output-format csv -d D/fetch -o b -- python3 bench.py --steps 1 \\
        --warmup 0 --no-plain --no-cpu --no-serve 2> D/fetch.log
    rocprofv3 --pmc WRITE_SIZE --output-format csv -d D/write -o b -- python3 bench.py --steps 1 \\
        --warmup 0 --no-plain --no-cpu --no-serve
    python3 -m measure.pmc_traffic --dir D --out profiles/pmc_traffic_mix1024_fp32.json

Each counter reads KiB per dispatch.  The copy kernel moves exactly `--mib`
MiB each way, so counter / true bytes there is the factor that turns the
counter into bytes.  A step launch of bench.py's kernel statistics is one
step of the whole batch: for PRIA a kernel of each iterate precision, for
PRAQ one kernel; its traffic is the sum over all dispatches of the kernel's
instantiations divided by the number of steps (the dispatch count of the
most-dispatched instantiation).  The algorithmic bytes per launch are those
bench.py printed in the FETCH_SIZE pass (its timed-launch solve).
"""
from __future__ import annotations

import argparse
import collections
import csv
import glob
import json
import re
from pathlib import Path

KERNELS = ("wpria_yz", "wpria_x", "wpraq_edge", "wpraq_node")


def counter_rows(d: Path):
    for f in glob.glob(str(d / "**" / "*counter_collection.csv"), recursive=True):
        yield from csv.DictReader(open(f))


def per_dispatch(d: Path):
    """{(kernel name, dispatch id): KiB}, summed over the counter's instances."""
    out = collections.defaultdict(float)
    for r in counter_rows(d):
        out[(r["Kernel_Name"], r["Dispatch_Id"])] += float(r["Counter_Value"])
    return out


def factor(d: Path, mib: int) -> float:
    """Counter KiB per true KiB, from the copy dispatches (the largest ones)."""
    vals = [v for (name, _), v in per_dispatch(d).items() if "copy" in name.lower()]
    big = max(vals)
    vals = [v for v in vals if v > 0.5 * big]
    return (sum(vals) / len(vals)) / (mib * 1024.0)


def step_traffic(d: Path, kernel: str, fac: float) -> float:
    """True bytes per step launch of `kernel`."""
    total, count = 0.0, collections.Counter()
    for (name, _), v in per_dispatch(d).items():
        m = re.search(r"\b" + kernel + r"\b(<[^>(]*>)?", name)
        if not m:
            continue
        total += v
        count[m.group(0)] += 1
    steps = max(count.values())
    return total / fac * 1024.0 / steps


def algorithmic(log: Path) -> dict:
    """Bytes per launch from bench.py's lines
    'kernel K: N launches, U us avg, G GB/s algorithmic'."""
    out = {}
    for line in log.read_text().splitlines():
        m = re.match(r"kernel (\w+): (\d+) launches, ([\d.]+) us avg, ([\d.]+) GB/s", line)
        if m:
            out[m.group(1)] = float(m.group(3)) * float(m.group(4)) * 1e3
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--dir", required=True)
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--out", required=True)
    args = ap.parse_args()
    d = Path(args.dir)
    ff = factor(d / "copy_fetch", args.mib)
    wf = factor(d / "copy_write", args.mib)
    alg = algorithmic(d / "fetch.log")
    rec = dict(workload="batched mixture of 1024 instances, FP32 iterates, tolerance 1e-6",
               fetch_factor=ff, write_factor=wf, kernels={})
    for k in KERNELS:
        fb = step_traffic(d / "fetch", k, ff)
        wb = step_traffic(d / "write", k, wf)
        rec["kernels"][k] = dict(fetch_bytes=fb, write_bytes=wb, traffic_bytes=fb + wb,
                                 algorithmic_bytes=alg.get(k),
                                 ratio=(fb + wb) / alg[k] if alg.get(k) else None)
    Path(args.out).write_text(json.dumps(rec, indent=1) + "\n")
    print(json.dumps(rec, indent=1))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
