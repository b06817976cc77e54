"""Writes c2_roof_dual.json: the PRAQ relaxation value of the large Chung-Lu
instance (n = 1e6, m = 1.6e7, Gaussian weights; workloads.praq_chung_lu(0))
by the oracle's int64 max-flow roof dual (capacities on the 2^-30 grid).

    python tests/golden/make_c2_roof_dual.py      (about 40 s, 3 GB of RAM)
"""
import json
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from reformulation_amd import workloads  # noqa: E402

SCALE = 2.0 ** 30


def main() -> None:
    w = workloads.praq_chung_lu(0)
    t0 = time.perf_counter()
    value = oracle.praq_roof_dual(w.n, w.diag, w.rows, w.cols, w.vals, scale=SCALE)
    seconds = time.perf_counter() - t0
    # each arc's capacity is rounded by at most 1/2 on the grid of a / 2 *
    # 2 scale: the cut moves by at most (arcs) / (4 scale)
    arcs = 2 * (len(w.vals) + w.n)
    out = dict(instance="workloads.praq_chung_lu(0)", n=int(w.n), m=int(len(w.vals)),
               roof_dual=value, scale=SCALE, error_bound=arcs / (4.0 * SCALE),
               seconds=round(seconds, 1))
    path = Path(__file__).resolve().parent / "c2_roof_dual.json"
    path.write_text(json.dumps(out, indent=1) + "\n")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
