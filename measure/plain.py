"""Plain one-instance solves of the two single-instance configurations, GPU
against the host oracle on the same instances.

    python -m measure.plain [--seeds 10] [--cpu]
    python -m measure.plain --family pria --items 350 --periods 450 \
        --tolerance 1e-4 --precisions fp32,fp64 --seeds 1

  praq   Erdos-Renyi n = 2000, m = 8000, integer weights (`workloads.praq_er`),
         FP64 iterates, 1e-6 relative gap, seeds 0..
  pria   N = 50 items, T = 200 periods (`workloads.pria_c1`), FP64 iterates,
         1e-6 relative gap, seeds 0..

Other sizes, tolerances and iterate precisions by --items/--periods,
--tolerance and --precisions (fp64, fp32: FP32 iterates with the FP64
handover).  Per instance: the GPU solve's wall time through the C ABI (`praq_solve_ex` /
`pria_solve_ex`: host build, upload, solve, download), its iterations and
bounds; with --cpu also the oracle's (FP64 C restatement, OpenMP over the
host's cores) wall time, iterations and |U_gpu - U_oracle| / (1 + |U|).  One
JSON line per family and precision with the per-instance records and medians.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np


def _praq(seeds: int, cpu: bool, tol: float = 1e-6, f32: bool = False):
    from reformulation_amd import praq, workloads
    out = []
    for seed in range(seeds):
        inst = workloads.praq_er(seed, 2000, 8000)
        problem = inst.problem()
        params = praq.Parameters.with_auto_step(-1, tol, use_float32=f32)
        praq.solve_with_info(problem, params)  # first call pays the context
        t0 = time.perf_counter()
        x, _, info = praq.solve_with_info(problem, params)
        rec = dict(seed=seed, seconds=time.perf_counter() - t0, iterations=int(info.iterations),
                   upper=info.upper, lower=info.lower, tier=int(info.tier))
        if cpu:
            import oracle
            t0 = time.perf_counter()
            o = oracle.praq_solve(inst.n, inst.diag, inst.rows, inst.cols, inst.vals,
                                  oracle.praq_params(tolerance=tol))
            rec.update(cpu_seconds=time.perf_counter() - t0, cpu_iterations=int(o.iterations),
                       upper_diff=abs(info.upper - o.upper) / (1 + abs(o.upper)))
        print(f"praq seed {seed}: {rec}", file=sys.stderr, flush=True)
        out.append(rec)
    return out


def _pria(seeds: int, cpu: bool, tol: float = 1e-6, f32: bool = False, items: int = 50,
          periods: int = 200):
    from reformulation_amd import pria, workloads
    out = []
    for seed in range(seeds):
        inst = workloads.pria_c1(seed, items, periods)
        problem = inst.problem()
        params = pria.Parameters.with_auto_step(-1, tol, use_float32=f32)
        t0 = time.perf_counter()
        x, y, z, info = pria.solve_with_info(problem, params)
        rec = dict(seed=seed, seconds=time.perf_counter() - t0, iterations=int(info.iterations),
                   upper=info.upper, lower=info.lower, tier=int(info.tier))
        if cpu:
            import oracle
            t0 = time.perf_counter()
            o = oracle.pria_solve(inst.order_costs, inst.joint_order_cost, inst.holding_costs,
                                  oracle.pria_params(tolerance=tol))
            rec.update(cpu_seconds=time.perf_counter() - t0, cpu_iterations=int(o.iterations),
                       upper_diff=abs(info.upper - o.upper) / (1 + abs(o.upper)))
        print(f"pria seed {seed}: {rec}", file=sys.stderr, flush=True)
        out.append(rec)
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--seeds", type=int, default=10)
    ap.add_argument("--cpu", action="store_true", help="also time the oracle")
    ap.add_argument("--family", choices=("praq", "pria", "both"), default="both")
    ap.add_argument("--tolerance", type=float, default=1e-6)
    ap.add_argument("--precisions", default="fp64")
    ap.add_argument("--items", type=int, default=50)
    ap.add_argument("--periods", type=int, default=200)
    args = ap.parse_args()
    for family in ("praq", "pria"):
        if args.family not in (family, "both"):
            continue
        for prec in args.precisions.split(","):
            f32 = prec == "fp32"
            if family == "praq":
                recs = _praq(args.seeds, args.cpu, args.tolerance, f32)
            else:
                recs = _pria(args.seeds, args.cpu, args.tolerance, f32, args.items, args.periods)
            _report(args, family, prec, recs)
    return 0


def _report(args, family, prec, recs):
    summary = dict(family=family, seeds=args.seeds, precision=prec, tolerance=args.tolerance,
                   median_seconds=float(np.median([r["seconds"] for r in recs])),
                   median_iterations=float(np.median([r["iterations"] for r in recs])))
    if args.cpu:
        summary.update(median_cpu_seconds=float(np.median([r["cpu_seconds"] for r in recs])),
                       cpu_cores=int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                       max_upper_diff=float(max(r["upper_diff"] for r in recs)))
    if family == "pria":
        summary.update(items=args.items, periods=args.periods)
    print(json.dumps(dict(summary, instances=recs)), flush=True)

if __name__ == "__main__":
    sys.exit(main())
