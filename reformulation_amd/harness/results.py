"""The reference harness's result records, field for field.

QUBO rows: qubo/harness/results.py:24-80.  JRP rows (items/periods/entries
instead of n/edges): jrp/harness/results.py:25-70.  `result_class` prefers the
reference harness's own classes when `harness.results` is importable.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

STATUSES = ("optimal", "time_limit", "iteration_limit", "error", "unavailable")


@dataclass
class QuboResult:
    backend: str
    instance: str
    family: str
    n: int
    edges: int
    status: str
    objective: float = math.nan
    lower: float = math.nan
    upper: float = math.nan
    iterations: int = -1
    setup_seconds: float = math.nan
    solve_seconds: float = math.nan
    optimality_tolerance: float = math.nan
    feasibility_tolerance: float = math.nan
    peak_host_bytes: float = math.nan
    peak_device_bytes: float = math.nan
    note: str = ""

    def __post_init__(self):
        if self.status not in STATUSES:
            raise ValueError(f"unknown status {self.status!r}, expected one of {STATUSES}")


@dataclass
class JrpResult:
    backend: str
    instance: str
    family: str
    items: int
    periods: int
    entries: int
    status: str
    objective: float = math.nan
    lower: float = math.nan
    upper: float = math.nan
    iterations: int = -1
    setup_seconds: float = math.nan
    solve_seconds: float = math.nan
    optimality_tolerance: float = math.nan
    feasibility_tolerance: float = math.nan
    peak_host_bytes: float = math.nan
    peak_device_bytes: float = math.nan
    note: str = ""

    def __post_init__(self):
        if self.status not in STATUSES:
            raise ValueError(f"unknown status {self.status!r}, expected one of {STATUSES}")


def result_class(kind: str):
    """The harness's own `Result` when importable, else the local twin."""
    try:
        from harness.results import Result  # the reference harness, when on sys.path
        fields = getattr(Result, "__dataclass_fields__", {})
        if kind == "qubo" and "edges" in fields:
            return Result
        if kind == "jrp" and "entries" in fields:
            return Result
    except Exception:
        pass
    return QuboResult if kind == "qubo" else JrpResult


def memory_fields(info, sample_memory: bool) -> dict:
    """peak_device_bytes: the device memory the solve held at its peak, as the
    library counts it (its images and work areas); peak_host_bytes: this
    process's resident-set high-water mark (the solve runs in-process, so it
    includes everything the process did before).  Only with sample_memory,
    as the reference's backends sample only then."""
    if not sample_memory:
        return {}
    import resource
    return dict(peak_device_bytes=float(info.device_bytes),
                peak_host_bytes=float(resource.getrusage(resource.RUSAGE_SELF).ru_maxrss) * 1024.0)
