"""Continuous serving of PRAQ and PRIA solves (include/reform_serve.h).

    srv = Server(use_float32=True)
    t = srv.submit_praq(problem, praq.Parameters.with_auto_step(50_000, 1e-6))
    srv.wait_idle()
    x, y, info = srv.praq_result(t)

A server owns a device memory pool (the HBM budget), host builder threads and
one serving thread.  Requests are built as they arrive, admitted in
submission order as their images fit in the pool, join the rounds of the
batched engine with their own Halpern state and precision, and retire as
they converge; their memory and state slots go to the next admissions.

`open_loop` drives a server with Poisson arrivals on its own clock and
returns each request's latency (completion - arrival); `fifo_latencies` is
the same request sequence through a one-at-a-time solver with measured
per-request service times (the plain solver's queue).
"""
from __future__ import annotations

import ctypes
import threading
import time
from dataclasses import dataclass

import numpy as np

from . import _native, praq as _praq, pria as _pria
from .batch import praq_native_params

_i64 = ctypes.c_int64
_dbl = ctypes.c_double
_int = ctypes.c_int


class ServerConfig(ctypes.Structure):
    _fields_ = [("use_float32", _int), ("host_threads", _int), ("hbm_budget", _i64),
                ("max_in_flight", _i64), ("borrow_inputs", _int), ("build_ahead", _int)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("ticket", _i64), ("family", _int), ("status", _int), ("iterations", _i64),
                ("restarts", _i64), ("switch_iteration", _i64), ("lower", _dbl),
                ("upper", _dbl), ("relative_gap", _dbl), ("submitted", _dbl),
                ("admitted", _dbl), ("completed", _dbl), ("device_bytes", _i64),
                ("bracket", _dbl)]


class ServerCounters(ctypes.Structure):
    _fields_ = [("submitted", _i64), ("admitted", _i64), ("completed", _i64),
                ("rejected", _i64), ("rounds", _i64), ("rebuilds", _i64), ("in_flight", _i64),
                ("bytes_in_use", _i64), ("bytes_peak", _i64), ("bytes_budget", _i64),
                ("retire_seconds", ctypes.c_double), ("rebuild_seconds", ctypes.c_double),
                ("in_flight_peak", _i64), ("handovers", _i64), ("handover_waits", _i64),
                ("cancelled", _i64)]


def _declare(L):
    if getattr(L, "_serve_declared", False):
        return
    vp = ctypes.c_void_p
    L.reform_server_default_config.restype = ServerConfig
    L.reform_server_default_config.argtypes = []
    L.reform_server_create.restype = vp
    L.reform_server_create.argtypes = [ctypes.POINTER(ServerConfig)]
    L.reform_server_destroy.restype = None
    L.reform_server_destroy.argtypes = [vp]
    L.reform_server_submit_praq.restype = _i64
    L.reform_server_submit_praq.argtypes = [vp, ctypes.POINTER(_native.PraqInstance),
                                            ctypes.POINTER(_native.PraqParameters),
                                            ctypes.POINTER(_native.PraqOptions)]
    L.reform_server_submit_pria.restype = _i64
    L.reform_server_submit_pria.argtypes = [vp, _i64, _i64, _native._dp, _dbl, _native._dp,
                                            ctypes.POINTER(_native.PriaParams),
                                            ctypes.POINTER(_native.PriaOptions)]
    L.reform_server_poll.restype = _i64
    L.reform_server_poll.argtypes = [vp, _native._ip, _i64]
    L.reform_server_wait_idle.restype = _int
    L.reform_server_wait_idle.argtypes = [vp, _dbl]
    L.reform_server_info.restype = _int
    L.reform_server_info.argtypes = [vp, _i64, ctypes.POINTER(RequestInfo)]
    L.reform_server_praq_result.restype = _int
    L.reform_server_praq_result.argtypes = [vp, _i64, _native._dp, _native._dp]
    L.reform_server_pria_result.restype = _int
    L.reform_server_pria_result.argtypes = [vp, _i64, _native._dp, _native._dp, _native._dp]
    L.reform_server_cancel.restype = _int
    L.reform_server_cancel.argtypes = [vp, _i64]
    L.reform_server_release.restype = _int
    L.reform_server_release.argtypes = [vp, _i64]
    L.reform_server_stats.restype = _int
    L.reform_server_stats.argtypes = [vp, ctypes.POINTER(ServerCounters)]
    L._serve_declared = True


@dataclass
class Info:
    ticket: int
    family: str
    status: int
    iterations: int
    restarts: int
    switch_iteration: int
    lower: float
    upper: float
    relative_gap: float
    submitted: float
    admitted: float
    completed: float
    device_bytes: int
    bracket: float

    @property
    def latency(self) -> float:
        return self.completed - self.submitted


class Server:
    """A running server; `borrow_inputs` keeps references to the submitted
    arrays (they must stay unchanged until the request finishes) instead of
    copying them."""

    def __init__(self, use_float32: bool = True, host_threads: int = 4,
                 hbm_budget: int = 200 * 10**9, max_in_flight: int = 16384,
                 borrow_inputs: bool = False, build_ahead: int = 32):
        self._L = _native.lib()
        _declare(self._L)
        c = self._L.reform_server_default_config()
        c.use_float32 = int(use_float32)
        c.host_threads = int(host_threads)
        c.hbm_budget = int(hbm_budget)
        c.max_in_flight = int(max_in_flight)
        c.borrow_inputs = int(borrow_inputs)
        c.build_ahead = int(build_ahead)
        self._h = self._L.reform_server_create(ctypes.byref(c))
        if not self._h:
            raise RuntimeError(f"reform_server_create failed: {_native.last_error()}")
        self.borrow_inputs = borrow_inputs
        self._shapes: dict[int, tuple] = {}
        self._keep: dict[int, tuple] = {}   # borrowed arrays, alive until release

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.reform_server_destroy(h)
            self._h = None
            self._keep.clear()

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit_praq(self, problem: _praq.QuboProblem, params: _praq.Parameters) -> int:
        p, o = praq_native_params(params)
        dense = problem.dense if params.dense_degree is None else params.dense_degree
        o.dense_degree = int(bool(dense))
        arrays = (problem.diag, problem.rows, problem.cols, problem.vals)
        inst = _native.PraqInstance(problem.n(), problem.m(), _native.dptr(arrays[0]),
                                    _native.iptr(arrays[1]), _native.iptr(arrays[2]),
                                    _native.dptr(arrays[3]))
        t = self._L.reform_server_submit_praq(self._h, ctypes.byref(inst), ctypes.byref(p),
                                              ctypes.byref(o))
        if t < 0:
            raise ValueError(f"reform_server_submit_praq failed ({t})")
        self._shapes[int(t)] = ("praq", problem.n(), problem.m())
        if self.borrow_inputs:
            self._keep[int(t)] = arrays
        return int(t)

    def submit_pria(self, problem: _pria.JrpProblem, params: _pria.Parameters) -> int:
        p = params.native()
        o = self._L.pria_default_options()
        o.tier = params.tier
        o.time_limit = params.time_limit
        k = np.ascontiguousarray(problem.order_costs, dtype=np.float64)
        h = np.ascontiguousarray(problem.holding_costs, dtype=np.float64)
        t = self._L.reform_server_submit_pria(self._h, problem.n(), problem.t(), _native.dptr(k),
                                              float(problem.joint_order_cost), _native.dptr(h),
                                              ctypes.byref(p), ctypes.byref(o))
        if t < 0:
            raise ValueError(f"reform_server_submit_pria failed ({t})")
        self._shapes[int(t)] = ("pria", problem.n(), problem.t())
        if self.borrow_inputs:
            self._keep[int(t)] = (k, h)
        return int(t)

    def poll(self, max_count: int = 4096) -> list[int]:
        buf = (ctypes.c_int64 * max_count)()
        k = self._L.reform_server_poll(self._h, ctypes.cast(buf, _native._ip), max_count)
        if k < 0:
            raise RuntimeError(f"reform_server_poll failed ({k})")
        return [int(buf[q]) for q in range(k)]

    def wait_idle(self, timeout: float = 0.0) -> None:
        st = self._L.reform_server_wait_idle(self._h, float(timeout))
        if st == -5:
            raise TimeoutError("requests still in flight")
        if st < 0:
            raise RuntimeError(f"serving failed ({st}): {_native.last_error()}")

    def info(self, ticket: int) -> Info:
        r = RequestInfo()
        st = self._L.reform_server_info(self._h, ticket, ctypes.byref(r))
        if st < 0:
            raise KeyError(ticket)
        return Info(int(r.ticket), "praq" if r.family == 0 else "pria", int(r.status),
                    int(r.iterations), int(r.restarts), int(r.switch_iteration), r.lower, r.upper,
                    r.relative_gap, r.submitted, r.admitted, r.completed, int(r.device_bytes),
                    r.bracket)

    def praq_result(self, ticket: int):
        _, n, m = self._shapes[ticket]
        x = np.zeros(n)
        y = np.zeros((max(m, 1), 2))
        st = self._L.reform_server_praq_result(self._h, ticket, _native.dptr(x), _native.dptr(y))
        if st < 0:
            raise RuntimeError(f"reform_server_praq_result failed ({st})")
        return x, y[:m], self.info(ticket)

    def pria_result(self, ticket: int):
        _, n, t = self._shapes[ticket]
        x, y, z = (np.zeros((n, t, t)) for _ in range(3))
        st = self._L.reform_server_pria_result(self._h, ticket, _native.dptr(x), _native.dptr(y),
                                               _native.dptr(z))
        if st < 0:
            raise RuntimeError(f"reform_server_pria_result failed ({st})")
        return x, y, z, self.info(ticket)

    def cancel(self, ticket: int) -> bool:
        """Stop an unfinished request: status -6 (no answer) if it had not
        been admitted, else 3 at its point at the next round.  False if the
        ticket had already finished (or was cancelled before)."""
        return self._L.reform_server_cancel(self._h, ticket) == 0

    def release(self, ticket: int) -> None:
        st = self._L.reform_server_release(self._h, ticket)
        if st < 0:
            raise KeyError(ticket)
        self._shapes.pop(ticket, None)
        self._keep.pop(ticket, None)

    def stats(self) -> dict:
        c = ServerCounters()
        st = self._L.reform_server_stats(self._h, ctypes.byref(c))
        if st < 0:
            raise RuntimeError(f"serving failed ({st}): {_native.last_error()}")
        return {name: (float if kind is ctypes.c_double else int)(getattr(c, name))
                for name, kind in ServerCounters._fields_}


def poisson_arrivals(rate: float, seconds: float, seed: int) -> np.ndarray:
    """Arrival times in [0, seconds) of a Poisson process of `rate` per second."""
    rng = np.random.default_rng(seed)
    count = int(rate * seconds * 1.2 + 64)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    while t[-1] < seconds:
        t = np.concatenate([t, t[-1] + np.cumsum(rng.exponential(1.0 / rate, count))])
    return t[t < seconds]


def open_loop(server: Server, arrivals: np.ndarray, submit, drain_timeout: float = 1800.0,
              on_done=None, progress=None, every: float = 10.0):
    """Submit request q at arrivals[q] seconds (submit(q) -> ticket) from this
    thread while a second thread collects completions; returns per request
    (arrival, completion) on the server clock plus the run's wall span.

    Arrivals do not wait for completions (open loop): a request late against
    its arrival time (the submitting thread fell behind) is submitted at
    once and its latency still counts from its scheduled arrival.
    `progress(completed, total)` is called every `every` seconds."""
    n = len(arrivals)
    tickets = np.full(n, -1, dtype=np.int64)
    arrive = np.zeros(n)
    done = np.full(n, np.nan)
    status = np.zeros(n, dtype=np.int64)
    by_ticket: dict[int, int] = {}
    lock = threading.Lock()
    finished = threading.Event()
    errors: list[BaseException] = []

    def collect():
        got = 0
        next_report = time.perf_counter() + every
        try:
            while got < n:
                if progress is not None and time.perf_counter() >= next_report:
                    next_report += every
                    progress(got, n)
                ready = server.poll()
                if not ready:
                    if finished.wait(0.001) and got >= n:
                        break
                    continue
                for t in ready:
                    with lock:
                        q = by_ticket.pop(t, None)
                    while q is None:   # submitted, not yet recorded by the submitter
                        time.sleep(0.0001)
                        with lock:
                            q = by_ticket.pop(t, None)
                    inf = server.info(t)
                    done[q] = inf.completed
                    status[q] = inf.status
                    if on_done is not None:
                        on_done(q, t)
                    server.release(t)
                    got += 1
        except BaseException as e:   # reported by the caller
            errors.append(e)
        finally:
            finished.set()

    worker = threading.Thread(target=collect, daemon=True)
    worker.start()
    t0 = time.perf_counter()
    offset = None
    for q in range(n):
        wait = arrivals[q] - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        t = submit(q)
        with lock:
            by_ticket[t] = q
        if offset is None:
            # server clock at this process's t0: submission stamps minus elapsed
            offset = server.info(t).submitted - (time.perf_counter() - t0)
        tickets[q] = t
        arrive[q] = offset + arrivals[q]
    worker.join(drain_timeout)
    finished.set()
    if errors:
        raise errors[0]
    if worker.is_alive():
        raise TimeoutError("requests still in flight after the drain timeout")
    return arrive, done, status


def fifo_latencies(arrivals: np.ndarray, service: np.ndarray) -> np.ndarray:
    """Completion - arrival for a single server taking requests in arrival
    order, request q busy for service[q] seconds."""
    free = 0.0
    out = np.empty(len(arrivals))
    for q, (a, s) in enumerate(zip(arrivals, service)):
        start = max(a, free)
        free = start + s
        out[q] = free - a
    return out


def summarize(arrive: np.ndarray, done: np.ndarray) -> dict:
    """p50 / p99 latency (ms) and completions per second over the span from
    the first arrival to the last completion."""
    lat = done - arrive
    span = float(np.nanmax(done) - np.min(arrive)) if len(lat) else 0.0
    return {"requests": int(len(lat)),
            "p50_ms": float(np.percentile(lat, 50) * 1e3) if len(lat) else None,
            "p99_ms": float(np.percentile(lat, 99) * 1e3) if len(lat) else None,
            "solves_per_s": float(len(lat) / span) if span > 0 else None}
