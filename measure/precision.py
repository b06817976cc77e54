"""Precision study of one large PRAQ instance: FP32 iterates (FP64
reductions, FP64 finish) against all-FP64 iterates, at each tolerance.

    python -m measure.precision [--n 1000000] [--tolerances 1e-4,1e-6]

The instance is the Chung-Lu power-law graph of `workloads.praq_chung_lu`
(exponent 2.5, mean degree 32, q_ij and c_i ~ N(0, 1); n = 1e6 gives about
1.6e7 edges).  Each solve runs the plain solver (`praq_solve_ex`, device tier)
from the centroid; the line per tolerance reports both runs' iterations,
restarts, bounds, gap and wall time, and the FP32 run's objective error
against the FP64 run, |U32 - U64| / (1 + |U64|), the same for L, and the
largest primal difference.  One JSON line per tolerance on stdout.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import sys
import time

import numpy as np


def study(n: int = 1_000_000, seed: int = 0, degree: float = 32.0,
          tolerances=(1e-4, 1e-6), max_iter: int = 200_000, time_limit: float = 600.0,
          log=lambda text: print(text, file=sys.stderr, flush=True)) -> list[dict]:
    """One record per tolerance (see the module text)."""
    from reformulation_amd import praq, workloads
    t0 = time.perf_counter()
    inst = workloads.praq_chung_lu(seed, n, mean_degree=degree)
    problem = inst.problem()
    log(f"instance {inst.name}: n={inst.n} m={inst.edges} built in "
        f"{time.perf_counter() - t0:.1f} s")
    out = []
    for tol in tolerances:
        runs = {}
        for f32 in (False, True):
            params = praq.Parameters.with_auto_step(max_iter, tol)
            params = dataclasses.replace(params, use_float32=f32, time_limit=time_limit)
            t1 = time.perf_counter()
            x, _, info = praq.solve_with_info(problem, params)
            dt = time.perf_counter() - t1
            runs[f32] = (x, info, dt)
            log(f"tol {tol:g} {'fp32' if f32 else 'fp64'}: {info.iterations} iterations, "
                f"gap {info.relative_gap:.2e}, {dt:.1f} s")
        x64, i64, t64 = runs[False]
        x32, i32, t32 = runs[True]

        def rec(info, dt):
            return dict(iterations=int(info.iterations), restarts=int(info.restarts),
                        status=int(info.status), lower=info.lower, upper=info.upper,
                        gap=info.relative_gap, bracket=info.bracket, seconds=dt,
                        switch_iteration=int(info.switch_iteration))

        scale = 1.0 + abs(i64.upper)
        out.append(dict(
            instance=inst.name, n=inst.n, m=inst.edges, tolerance=tol,
            fp64=rec(i64, t64), fp32=rec(i32, t32),
            upper_error=abs(i32.upper - i64.upper) / scale,
            lower_error=abs(i32.lower - i64.lower) / (1.0 + abs(i64.lower)),
            primal_max_diff=float(np.max(np.abs(x32 - x64)))))
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--degree", type=float, default=32.0)
    ap.add_argument("--tolerances", default="1e-4,1e-6")
    ap.add_argument("--max-iter", type=int, default=200_000)
    ap.add_argument("--time-limit", type=float, default=600.0)
    args = ap.parse_args()
    tols = [float(t) for t in args.tolerances.split(",")]
    for rec in study(args.n, args.seed, args.degree, tols, args.max_iter, args.time_limit):
        print(json.dumps(rec), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
