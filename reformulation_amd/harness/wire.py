"""The file CLIs' wire formats (reformulation_amd/tools/*_solve_file.cpp),
which are the reference CLIs' (qubo/praq-cuda/tools/praq_solve_file.cu:7-17,
jrp/pria-cuda/tools/pria_solve_file.cu:7-17): little-endian int64 / float64
records, read and written with numpy."""
from __future__ import annotations

from pathlib import Path

import numpy as np

PRAQ_MAGIC = 0x51415250
PRIA_MAGIC = 0x41495250
BIN = Path(__file__).resolve().parents[1] / "bin"


def write_praq_instance(path, n, diag, rows, cols, vals) -> None:
    with open(path, "wb") as f:
        np.array([PRAQ_MAGIC, n, len(rows)], dtype="<i8").tofile(f)
        np.asarray(diag, dtype="<f8").tofile(f)
        np.asarray(rows, dtype="<i8").tofile(f)
        np.asarray(cols, dtype="<i8").tofile(f)
        np.asarray(vals, dtype="<f8").tofile(f)


def read_praq_result(path) -> dict:
    raw = Path(path).read_bytes()
    n, iterations, status = np.frombuffer(raw[:24], dtype="<i8")
    lower, upper = np.frombuffer(raw[24:40], dtype="<f8")
    x = np.frombuffer(raw[40:40 + 8 * int(n)], dtype="<f8").copy()
    return dict(n=int(n), iterations=int(iterations), status=int(status), lower=float(lower),
                upper=float(upper), x=x)


def write_pria_instance(path, order_costs, joint_order_cost, holding_costs) -> None:
    h = np.asarray(holding_costs, dtype="<f8")
    n, t = h.shape[0], h.shape[1]
    with open(path, "wb") as f:
        np.array([PRIA_MAGIC, n, t], dtype="<i8").tofile(f)
        np.array([joint_order_cost], dtype="<f8").tofile(f)
        np.asarray(order_costs, dtype="<f8").tofile(f)
        h.tofile(f)


def read_pria_result(path, with_solution: bool = False) -> dict:
    raw = Path(path).read_bytes()
    n, t, iterations, status = (int(v) for v in np.frombuffer(raw[:32], dtype="<i8"))
    lower, upper, gap = (float(v) for v in np.frombuffer(raw[32:56], dtype="<f8"))
    out = dict(n=n, t=t, iterations=iterations, status=status, lower=lower, upper=upper, gap=gap)
    if with_solution:
        size = n * t * t
        arr = np.frombuffer(raw[56:56 + 24 * size], dtype="<f8")
        out.update(x=arr[:size].reshape(n, t, t), y=arr[size:2 * size].reshape(n, t, t),
                   z=arr[2 * size:].reshape(n, t, t))
    return out


def parse_stdout(text: str) -> dict:
    """The CLIs' `key: value` lines."""
    out = {}
    for line in text.splitlines():
        if ": " in line:
            key, value = line.split(": ", 1)
            out[key.strip()] = value.strip()
    return out
