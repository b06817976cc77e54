"""Throughput of the batched PRAQ/PRIA solver on one MI355X.

Workload: the batched mixture -- 1024 instances, 70% PRAQ (Erdos-Renyi, n
log-uniform in [1e3, 1e5], mean degree U[4, 32], integer weights) and 30%
PRIA (N in {10..100}, T in {50..300}), instance k drawn with seed k -- solved
to 1e-6 relative duality gap with FP32 iterates and FP64 reductions.  A step
is one solve of the whole batch from inputs resident in HBM by the batched
mode (block tables over all instances, per-instance restart state on the
device, one check interval per hipGraph replay); `value` is solves/s over
the timed steps.  Beside it: the plain one-at-a-time GPU solver on the same
resident instances (solves/s); serving -- open-loop Poisson arrivals at
`--rates` requests/s for `--serve-seconds` each, request k an instance drawn
from the same mixture with its tolerance drawn from {1e-4, 1e-6}, through
the serving loop (HBM-budgeted admission, continuous batching, backfill),
p50/p99 latency (completion - arrival) and solves/s, beside the same
arrivals through a one-at-a-time queue whose service times are the plain
solver's measured per-instance times with inputs resident (by default 50
and 200 requests/s, below saturation -- `latency_ms` quotes the first --
and 1000 requests/s, above it: `served_saturated` is the served rate there
against the plain queue's); the
roofline of the dominant step kernel --
its launches' summed algorithmic bytes (PRAQ node n(4+7b) + 2m(4+b), edge
m(8+7b); PRIA x 7b ND, y/z 8b ND per step, b = 4 for FP32) over their summed
durations, each launch bracketed by HIP events on its stream in one more
solve of the same batch after the timed steps, with `traffic` the measured
HBM bytes per launch of that kernel (two rocprofv3 counter passes over this
workload, calibrated on a device copy, kept in profiles/); the oracle (FP64 C restatement, OpenMP) timed on a bounded sample
of the same mixture on the host's cores; and the two single-instance
configurations through the plain solver (PRAQ Erdos-Renyi n = 2000,
m = 8000; PRIA N = 50, T = 200; FP64 to 1e-6), GPU and oracle medians
(`single_instance`); and each family of the mixture batched on its own
(`families_alone`: its solve seconds and its step kernels' live rates with
the chip to themselves, beside the side-by-side step time).  `--precision` prints the FP32-against-FP64 study on
the large Chung-Lu PRAQ instance instead.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--count 1024]
    python bench.py --serve-only --rates 50,200,1000 --serve-seconds 120
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_HBM_GBS = 8000.0
TRAFFIC = ROOT / "profiles" / "pmc_traffic_mix1024_fp32.json"


def measured_traffic(kernel: str, count: int, use_f32: bool, tolerance: float):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE /
    WRITE_SIZE passes over this workload (copy-calibrated), when the run is
    that workload; else None."""
    if not (TRAFFIC.exists() and count == 1024 and use_f32 and tolerance == 1e-6):
        return None
    k = json.loads(TRAFFIC.read_text())["kernels"].get(kernel)
    return None if k is None else k["traffic_bytes"]


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def build_batch(members, tolerance: float, use_float32: bool, max_iter: int):
    from reformulation_amd import praq, pria
    from reformulation_amd.batch import Batch
    b = Batch(use_float32=use_float32)
    index = []
    for kind, inst in members:
        if kind == "praq":
            params = praq.Parameters.with_auto_step(max_iter, tolerance)
            index.append(("praq", b.add_praq(inst.problem(), params), inst))
        else:
            params = pria.Parameters.with_auto_step(max_iter, tolerance)
            index.append(("pria", b.add_pria(inst.problem(), params), inst))
    return b, index


def families_alone(members, tolerance: float, use_float32: bool, side_by_side_s: float) -> dict:
    """Each family of the mixture batched on its own (the other family's
    instances left out): solve seconds, and per step kernel of the family
    its live average and algorithmic rate -- the kernels with the chip to
    themselves, against the mixture's run where the two families' rounds
    share it."""
    out = dict(side_by_side_s=side_by_side_s)
    for kind in ("pria", "praq"):
        b, _ = build_batch([m for m in members if m[0] == kind], tolerance, use_float32, 200_000)
        b.upload()
        b.solve("batched")
        dt = min(b.solve("batched") for _ in range(2))
        b.set_timing(True)
        b.solve("batched")
        b.set_timing(False)
        ks = {k["name"]: dict(avg_launch_us=k["ms"] / k["launches"] * 1e3,
                              gbs=k["bytes"] / (k["ms"] * 1e-3) / 1e9,
                              frac=k["bytes"] / (k["ms"] * 1e-3) / 1e9 / PEAK_HBM_GBS)
              for k in b.kernel_stats() if k["launches"]}
        b.close()
        out[kind] = dict(solve_s=dt, kernels=ks)
        log(f"{kind} family alone: {dt:.3f} s per batched solve; "
            + ", ".join(f"{n} {v['avg_launch_us']:.0f} us {v['frac']:.1%}" for n, v in ks.items()))
    out["one_after_the_other_s"] = out["pria"]["solve_s"] + out["praq"]["solve_s"]
    return out


def cpu_baseline(members, tolerance: float, budget_s: float):
    """The oracle (FP64, all host cores) on the first instances of the mixture
    until `budget_s` of work."""
    import oracle
    cores = os.cpu_count() or 1
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    solved = 0
    t0 = time.perf_counter()
    for kind, inst in members:
        if kind == "praq":
            oracle.praq_solve(inst.n, inst.diag, inst.rows, inst.cols, inst.vals,
                              oracle.praq_params(tolerance=tolerance, max_iter=200_000))
        else:
            oracle.pria_solve(inst.order_costs, inst.joint_order_cost, inst.holding_costs,
                              oracle.pria_params(tolerance=tolerance, max_iter=200_000))
        solved += 1
        if time.perf_counter() - t0 > budget_s:
            break
    elapsed = time.perf_counter() - t0
    return dict(value=solved / elapsed, unit="solves/s", cores=int(os.environ["OMP_NUM_THREADS"]),
                kind="port",
                sample=f"first {solved} instances of the mixture, FP64 oracle, {elapsed:.1f} s")


def plain_service_times(index, tolerance_batch, count, solved=()):
    """Per mixture member: the plain solver's seconds at each tolerance, from
    sequential solves of resident batches (those of `solved` have just run
    one)."""
    out = {}
    for tol, b in tolerance_batch.items():
        if tol not in solved:
            b.solve("sequential")
        sp, sr = b.instance_seconds()
        out[tol] = np.array([sp[q] if kind == "praq" else sr[q] for kind, q, _ in index])
        log(f"plain service times at {tol:g}: mean {out[tol].mean() * 1e3:.1f} ms, "
            f"total {out[tol].sum():.1f} s over {count}")
    return out


def serving(members, rate: float, seconds: float, service, use_f32: bool, threads: int,
            seed: int = 2024):
    """One open-loop run of the serving loop and the plain queue on the same
    arrivals."""
    from reformulation_amd import praq, pria, serve
    rng = np.random.default_rng(seed)
    arrivals = serve.poisson_arrivals(rate, seconds, seed)
    picks = rng.integers(0, len(members), len(arrivals))
    tols = rng.choice([1e-4, 1e-6], len(arrivals))
    problems = [(kind, inst.problem()) for kind, inst in members]
    pp = {t: praq.Parameters.with_auto_step(200_000, t) for t in (1e-4, 1e-6)}
    rp = {t: pria.Parameters.with_auto_step(200_000, t) for t in (1e-4, 1e-6)}
    srv = serve.Server(use_float32=use_f32, host_threads=threads, borrow_inputs=True)

    def submit(q):
        kind, p = problems[picks[q]]
        t = tols[q]
        return srv.submit_praq(p, pp[t]) if kind == "praq" else srv.submit_pria(p, rp[t])

    t0 = time.perf_counter()
    arrive, done, status = serve.open_loop(
        srv, arrivals, submit, progress=lambda got, n: log(
            f"serving at {rate:g}/s: {got}/{n} done, {time.perf_counter() - t0:.0f} s, "
            f"{srv.stats()['in_flight']} in flight"))
    wall = time.perf_counter() - t0
    stats = srv.stats()
    srv.close()
    served = serve.summarize(arrive, done)
    served["converged"] = int(np.sum(status == 0))
    plain_lat = serve.fifo_latencies(arrivals, np.array(
        [service[t][m] for m, t in zip(picks, tols)]))
    plain = serve.summarize(arrivals, arrivals + plain_lat)
    out = dict(rate=rate, seconds=seconds, requests=len(arrivals), wall_s=wall,
               served=served, plain=plain,
               speedup=(served["solves_per_s"] / plain["solves_per_s"])
               if plain["solves_per_s"] else None,
               rounds=stats["rounds"], bytes_peak=stats["bytes_peak"],
               hbm_budget=stats["bytes_budget"], in_flight_peak=stats["in_flight_peak"],
               handovers=stats["handovers"], handover_waits=stats["handover_waits"],
               host_s=dict(retire=round(stats["retire_seconds"], 3),
                           rebuild=round(stats["rebuild_seconds"], 3)))
    log(f"serving at {rate:g}/s for {seconds:g} s: {json.dumps(out)}")
    return out


def precision_line() -> int:
    """The precision study (measure/precision.py) as one JSON line: `value`
    is the FP32-iterate run's objective error against the FP64-iterate run
    at tolerance 1e-6, |U32 - U64| / (1 + |U64|); each record carries both
    runs' iterations, handover iteration, bounds and certified brackets."""
    from measure.precision import study
    recs = study(log=log)
    at = next(r for r in recs if r["tolerance"] == 1e-6)
    line = dict(
        metric="FP32-iterate objective error vs FP64 iterates at 1e-6 (one large PRAQ)",
        value=at["upper_error"], unit="relative", n_gpus=1, higher_is_better=False,
        vs_baseline=None, dtype="f32", data="synthetic",
        config=dict(workload=f"PRAQ roof dual of {at['instance']} (Chung-Lu, exponent 2.5, "
                             f"mean degree 32, n {at['n']}, m {at['m']}), plain device tier, "
                             "FP32 iterates (FP64 reductions, FP64 finish) against FP64 "
                             "iterates at tolerances 1e-4 and 1e-6", parallelism="none"),
        records=recs)
    print(json.dumps(line), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--count", type=int, default=1024)
    ap.add_argument("--tolerance", type=float, default=1e-6)
    ap.add_argument("--fp64", action="store_true", help="FP64 iterates")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--rates", default="50,200,1000",
                    help="serving arrival rates, comma-separated (the highest saturates)")
    ap.add_argument("--serve-seconds", type=float, default=20.0)
    ap.add_argument("--serve-threads", type=int, default=12)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-only", action="store_true",
                    help="only the serving runs (one JSON line per rate)")
    ap.add_argument("--no-single", action="store_true",
                    help="skip the two single-instance plain configurations")
    ap.add_argument("--precision", action="store_true",
                    help="only the precision study: FP32 against FP64 iterates on the "
                         "Chung-Lu graph of 1e6 nodes (one JSON line)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps(dict(error="this build measures one GPU; --gpus must be 1 "
                                    "(instances are independent: N GPUs would each run "
                                    "their own batch)")))
        return 2

    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.precision:
        return precision_line()
    from reformulation_amd import workloads

    t0 = time.perf_counter()
    members = workloads.mix(args.count)
    log(f"generated {len(members)} instances in {time.perf_counter() - t0:.1f} s")
    use_f32 = not args.fp64
    batch, index = build_batch(members, args.tolerance, use_f32, 200_000)
    t0 = time.perf_counter()
    batch.upload()
    log(f"uploaded in {time.perf_counter() - t0:.1f} s, {batch.device_bytes / 1e9:.2f} GB on device")
    rates = [float(r) for r in args.rates.split(",") if r] if not args.no_serve else []

    def service_times(solved=()):
        other = 1e-4 if args.tolerance != 1e-4 else 1e-6
        b2, _ = build_batch(members, other, use_f32, 200_000)
        b2.upload()
        out = plain_service_times(index, {args.tolerance: batch, other: b2}, args.count, solved)
        b2.close()
        return out

    if args.serve_only:
        service = service_times()
        batch.close()   # the serving pool takes the HBM budget
        for rate in rates:
            out = serving(members, rate, args.serve_seconds, service, use_f32,
                          args.serve_threads)
            print(json.dumps(dict(workload="serving", count=args.count,
                                  dtype="f32" if use_f32 else "f64", **out)), flush=True)
        return 0

    for w in range(args.warmup):
        dt = batch.solve("batched")
        log(f"warmup {w}: {dt:.3f} s")
    torch.cuda.synchronize()
    times, kernels = [], {}
    for s in range(args.steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        batch.solve("batched")
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        km = batch.kernel_ms()
        log(f"step {s}: {times[-1]:.3f} s (praq family {km[0]:.1f} ms, pria family {km[1]:.1f} ms)")
    total = sum(times)
    value = args.count * args.steps / total

    # one more solve of the same batch with every step launch bracketed by
    # HIP events on its stream (rounds launched directly, not replayed)
    batch.set_timing(True)
    dt = batch.solve("batched")
    batch.set_timing(False)
    log(f"timed-launch solve: {dt:.3f} s")
    for k in batch.kernel_stats():
        kernels[k["name"]] = k

    infos = []
    for kind, q, inst in index:
        infos.append(batch.praq_result(q, False)[2] if kind == "praq" else batch.pria_result(q, False)[3])
    status = [i.status for i in infos]
    gaps = np.array([i.relative_gap for i in infos])
    iters = np.array([i.iterations for i in infos])
    log(f"converged {status.count(0)}/{len(status)}, max gap {gaps.max():.2e}, "
        f"iterations median {int(np.median(iters))} max {int(iters.max())}")

    dom = max(kernels, key=lambda name: kernels[name]["ms"])
    K = kernels[dom]
    achieved = K["bytes"] / (K["ms"] * 1e-3) / 1e9
    for name, k in sorted(kernels.items(), key=lambda kv: -kv[1]["ms"]):
        log(f"kernel {name}: {k['launches']} launches, {k['ms'] / k['launches'] * 1e3:.1f} us avg, "
            f"{k['bytes'] / (k['ms'] * 1e-3) / 1e9:.0f} GB/s algorithmic")

    alone = families_alone(members, args.tolerance, use_f32, min(times))

    plain = None
    if not args.no_plain:
        dt = batch.solve("sequential")
        plain = args.count / dt
        log(f"plain one-at-a-time: {dt:.2f} s -> {plain:.1f} solves/s")

    serve_runs = []
    if rates:
        service = service_times(() if args.no_plain else (args.tolerance,))
        batch.close()   # the serving pool takes the HBM budget
        for rate in rates:
            serve_runs.append(serving(members, rate, args.serve_seconds, service, use_f32,
                                      args.serve_threads))

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(members, args.tolerance, args.cpu_seconds)
        log(f"cpu baseline: {cpu}")
    single = None if args.no_single else single_instance(not args.no_cpu)

    ms = 1000.0 * total / args.steps
    line = dict(
        metric="solves/sec to 1e-6 rel KKT (served vs plain, 1 GPU); p50/p99 solve latency",
        value=value, unit="solves/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
        ms_per_step=ms, higher_is_better=True, scaling="weak", vs_baseline=None,
        dtype="f32" if use_f32 else "f64", data="synthetic",
        config=dict(workload=f"batched mixture of {args.count} instances (70% PRAQ ER n in "
                             "[1e3,1e5] deg U[4,32]; 30% PRIA N in [10,100], T in [50,300]) "
                             f"to {args.tolerance:g} relative gap, "
                             + ("FP32 iterates + FP64 reductions" if use_f32 else "FP64")
                             + ", batched mode (block tables, hipGraph rounds)",
                    count=args.count, tolerance=args.tolerance, parallelism="none"),
        plain_solves_per_s=plain, speedup_vs_plain=(value / plain) if plain else None,
        latency_ms=(dict(p50=serve_runs[0]["served"]["p50_ms"],
                         p99=serve_runs[0]["served"]["p99_ms"],
                         plain_p50=serve_runs[0]["plain"]["p50_ms"],
                         plain_p99=serve_runs[0]["plain"]["p99_ms"],
                         note=f"serving at {serve_runs[0]['rate']:g} requests/s")
                    if serve_runs else None),
        serving=serve_runs or None,
        served_saturated=(dict(rate=serve_runs[-1]["rate"],
                               solves_per_s=serve_runs[-1]["served"]["solves_per_s"],
                               plain_solves_per_s=serve_runs[-1]["plain"]["solves_per_s"],
                               ratio=serve_runs[-1]["speedup"])
                          if len(serve_runs) > 1 else None),
        converged=status.count(0), max_gap=float(gaps.max()),
        roofline=dict(bound="hbm", kernel=dom, achieved=achieved, peak=PEAK_HBM_GBS,
                      unit="GB/s", frac=achieved / PEAK_HBM_GBS,
                      traffic=measured_traffic(dom, args.count, use_f32, args.tolerance),
                      traffic_source=("profiles/pmc_traffic_mix1024_fp32.json: FETCH_SIZE and "
                                      "WRITE_SIZE per launch, copy-calibrated"),
                      launches=K["launches"], avg_launch_us=K["ms"] / K["launches"] * 1e3,
                      bytes_per_launch=K["bytes"] / K["launches"]),
        cpu_baseline=cpu,
        families_alone=alone,
        single_instance=single,
        tuning=tuning_in_use(),
    )
    print(json.dumps(line), flush=True)
    return 0


def single_instance(with_cpu: bool) -> dict:
    """The two single-instance configurations through the plain solver's C
    ABI (host build, upload, solve, download per call), FP64 iterates to
    1e-6: PRAQ Erdos-Renyi n = 2000, m = 8000 (seeds 0..4) and PRIA
    N = 50, T = 200 (seeds 0, 1); medians, with the oracle on the host's
    cores beside them (its iteration counts must agree)."""
    from measure import plain
    out = {}
    for key, workload, rows in (
            ("plain_praq_er2000", "PRAQ Erdos-Renyi n=2000 m=8000, FP64, 1e-6",
             plain._praq(5, with_cpu)),
            ("plain_pria_n50t200", "PRIA N=50 T=200, FP64, 1e-6", plain._pria(2, with_cpu))):
        rec = dict(workload=workload, instances=len(rows),
                   gpu_ms=1e3 * float(np.median([r["seconds"] for r in rows])),
                   iterations=[r["iterations"] for r in rows])
        if with_cpu:
            rec.update(cpu_ms=1e3 * float(np.median([r["cpu_seconds"] for r in rows])),
                       cpu_cores=int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                       iterations_equal=all(r["iterations"] == r["cpu_iterations"] for r in rows),
                       max_upper_diff=max(r["upper_diff"] for r in rows))
        out[key] = rec
        log(f"{key}: {rec}")
    return out


def tuning_in_use() -> dict:
    """Which shape table the batched tiers ran with: its file and the
    batched entries the library holds (unset entries keep the built-in
    choices)."""
    from reformulation_amd import autotune
    env = os.environ.get("REFORM_TUNING")
    source = "off" if env == "off" else (env or str(autotune.DEFAULT_TABLE.relative_to(ROOT)))
    entries = {k: v for k, v in autotune.table_in_use().items() if k.startswith("batch.") and v}
    return dict(table=source, batched_entries=entries)


if __name__ == "__main__":
    sys.exit(main())
