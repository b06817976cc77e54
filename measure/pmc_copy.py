"""Calibration for HBM byte counters: a plain device copy of known size.

    rocprofv3 --pmc FETCH_SIZE -- python3 measure/pmc_copy.py

copies a 1 GiB buffer into another `repeats` times; each copy kernel reads
and writes exactly 1 GiB, so FETCH_SIZE and WRITE_SIZE of those dispatches
(KiB) against 1048576 give the factor that turns the counters into bytes.
"""
from __future__ import annotations

import argparse

import torch


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--repeats", type=int, default=3)
    args = ap.parse_args()
    n = args.mib * (1 << 20) // 4
    a = torch.rand(n, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(args.repeats):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {args.mib} MiB {args.repeats} times", flush=True)


if __name__ == "__main__":
    main()
