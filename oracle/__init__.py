"""ctypes bindings of the oracle (oracle/*.c): an FP64 CPU restatement of the
reference's PRAQ and PRIA solves.

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the checker (or the timed CPU
baseline), never by the package itself.
"""
from __future__ import annotations

import ctypes
import math
import subprocess
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "build" / "liboracle.so"
REF_BIN = HERE / "_ref" / "pria_ref"
_i64 = ctypes.c_int64
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int64)


class OracleParams(ctypes.Structure):
    _fields_ = [("omega", ctypes.c_double), ("eta", ctypes.c_double),
                ("beta_sufficient", ctypes.c_double), ("beta_necessary", ctypes.c_double),
                ("beta_artificial", ctypes.c_double), ("theta_initial", ctypes.c_double),
                ("theta_target", ctypes.c_double), ("theta_decay", ctypes.c_double),
                ("tolerance", ctypes.c_double), ("max_iter", _i64), ("op_steps", _i64),
                ("warm_steps", _i64), ("use_halpern", _i64), ("dense_degree", _i64),
                ("interval_offset", ctypes.c_double), ("interval_scale", ctypes.c_double),
                ("interval_power", ctypes.c_double)]


class OracleInfo(ctypes.Structure):
    _fields_ = [("iterations", _i64), ("restarts", _i64), ("checks", _i64),
                ("lower", ctypes.c_double), ("upper", ctypes.c_double),
                ("upper_iterate", ctypes.c_double), ("relative_gap", ctypes.c_double),
                ("omega", ctypes.c_double), ("scale", ctypes.c_double), ("status", _i64)]


def build(force: bool = False) -> None:
    """Compile oracle/*.c (and oracle/_ref where the reference tree exists);
    `force` recompiles what is already built."""
    flags = ["-B"] if force else []
    subprocess.run(["make", "-s", *flags, "-C", str(HERE), "all"], check=True)
    if Path("/root/reference/jrp/pria-cuda/tests/ref_impl.hpp").exists():
        subprocess.run(["make", "-s", *flags, "-C", str(HERE), "ref"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = ctypes.CDLL(str(LIB))
        _lib.oracle_praq_solve.restype = ctypes.c_int
        _lib.oracle_pria_solve.restype = ctypes.c_int
        _lib.oracle_pria_primal_objective.restype = ctypes.c_double
        _lib.oracle_pria_dual_objective.restype = ctypes.c_double
        _lib.oracle_praq_check_interval.restype = _i64
        _lib.oracle_praq_check_interval.argtypes = [_i64, ctypes.c_double, ctypes.c_double,
                                                    ctypes.c_double]
        _lib.oracle_praq_roof_dual.restype = ctypes.c_double
        _lib.oracle_praq_roof_dual.argtypes = [_i64, _dp, _i64, ctypes.c_void_p,
                                               ctypes.c_void_p, _dp, ctypes.c_double,
                                               ctypes.POINTER(_i64)]
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _i(a):
    return a.ctypes.data_as(_ip) if a is not None else None


def praq_params(tolerance=1e-6, max_iter=-1, op_steps=16, dense_degree=False, omega=1.0,
                use_halpern=True, warm_steps=0, **kw) -> OracleParams:
    """praq-jax Parameters.with_auto_step defaults (solver.py:129-164)."""
    p = OracleParams(omega=omega, eta=0.999, beta_sufficient=math.exp(-1.0),
                     beta_necessary=0.5706, beta_artificial=0.2468, theta_initial=0.81930,
                     theta_target=0.42818, theta_decay=0.90474, tolerance=tolerance,
                     max_iter=-1 if max_iter is None else max_iter, op_steps=op_steps,
                     warm_steps=warm_steps, use_halpern=int(use_halpern),
                     dense_degree=int(dense_degree), interval_offset=0.0,
                     interval_scale=4.328085122666891, interval_power=0.0)
    for key, value in kw.items():
        setattr(p, key, value)
    return p


def pria_params(tolerance=1e-6, max_iter=-1, op_steps=16, omega=0.10835, use_halpern=True,
                warm_steps=0, **kw) -> OracleParams:
    """pria-jax Parameters.with_auto_step defaults (solver.py:150-188)."""
    p = OracleParams(omega=omega, eta=0.999 / math.sqrt(2.0),
                     beta_sufficient=math.exp(-1.0), beta_necessary=0.5706,
                     beta_artificial=0.2468, theta_initial=0.81930, theta_target=0.42818,
                     theta_decay=0.90474, tolerance=tolerance,
                     max_iter=-1 if max_iter is None else max_iter,
                     op_steps=max(op_steps, 2), warm_steps=warm_steps,
                     use_halpern=int(use_halpern), dense_degree=0)
    for key, value in kw.items():
        setattr(p, key, value)
    return p


@dataclass
class Solved:
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray | None
    iterations: int
    restarts: int
    lower: float
    upper: float
    upper_iterate: float
    relative_gap: float
    status: int
    scale: float
    omega: float


def praq_roof_dual(n, diag, rows, cols, vals, scale: float = 2.0 ** 30) -> float:
    """The PRAQ relaxation's value by max-flow (maxflow.c): capacities on
    the 1 / scale grid, exact for weights on it (integers with scale 1)."""
    diag = np.ascontiguousarray(diag, dtype=np.float64)
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int64)
    flow = _i64(0)
    v = lib().oracle_praq_roof_dual(_i64(n), _d(diag), _i64(len(vals)),
                                    rows.ctypes.data_as(ctypes.c_void_p),
                                    cols.ctypes.data_as(ctypes.c_void_p), _d(vals),
                                    float(scale), ctypes.byref(flow))
    if v != v:
        raise ValueError("oracle_praq_roof_dual failed")
    return float(v)


def praq_solve(n, diag, rows, cols, vals, params: OracleParams | None = None,
               warm=None) -> Solved:
    params = params or praq_params()
    diag = np.ascontiguousarray(diag, dtype=np.float64)
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int64)
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    m = vals.size
    x = np.zeros(n)
    y = np.zeros((max(m, 1), 2))
    wx = wy = None
    if warm is not None:
        wx = np.ascontiguousarray(warm[0], dtype=np.float64)
        wy = np.ascontiguousarray(warm[1], dtype=np.float64).reshape(-1)
    info = OracleInfo()
    st = lib().oracle_praq_solve(_i64(n), _i64(m), _d(diag), _i(rows), _i(cols), _d(vals),
                                 ctypes.byref(params), _d(wx), _d(wy), _d(x), _d(y),
                                 ctypes.byref(info))
    if st < 0:
        raise ValueError(f"oracle_praq_solve rejected the instance ({st})")
    return Solved(x, y[:m], None, info.iterations, info.restarts, info.lower, info.upper,
                  info.upper_iterate, info.relative_gap, info.status, info.scale, info.omega)


def praq_check_interval(n, offset=0.0, scale=4.328085122666891, power=0.0) -> int:
    return int(lib().oracle_praq_check_interval(n, offset, scale, power))


def praq_bounds(n, diag, rows, cols, vals, x, y):
    lo, up = ctypes.c_double(), ctypes.c_double()
    lib().oracle_praq_bounds(_i64(n), _i64(len(vals)),
                             _d(np.ascontiguousarray(diag, dtype=np.float64)),
                             _i(np.ascontiguousarray(rows, dtype=np.int64)),
                             _i(np.ascontiguousarray(cols, dtype=np.int64)),
                             _d(np.ascontiguousarray(vals, dtype=np.float64)),
                             _d(np.ascontiguousarray(x, dtype=np.float64)),
                             _d(np.ascontiguousarray(y, dtype=np.float64).reshape(-1)),
                             ctypes.byref(lo), ctypes.byref(up))
    return lo.value, up.value


def praq_structure(n, diag, rows, cols, vals):
    """(scaled weights w_e, covering flags, scale) of the sparse build."""
    m = len(vals)
    w = np.zeros(max(m, 1))
    pos = np.zeros(max(m, 1), dtype=np.uint8)
    sc = ctypes.c_double()
    st = lib().oracle_praq_structure(_i64(n), _i64(m), _d(np.ascontiguousarray(diag, dtype=np.float64)),
                                     _i(np.ascontiguousarray(rows, dtype=np.int64)),
                                     _i(np.ascontiguousarray(cols, dtype=np.int64)),
                                     _d(np.ascontiguousarray(vals, dtype=np.float64)), _d(w),
                                     pos.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                     ctypes.byref(sc))
    if st:
        raise ValueError(st)
    return w[:m], pos[:m].astype(bool), sc.value


def praq_raw_step(n, diag, rows, cols, vals, x1, y1, omega=1.0, eta=0.999, dense_degree=False):
    m = len(vals)
    x2 = np.zeros(n)
    y2 = np.zeros((max(m, 1), 2))
    lib().oracle_praq_raw_step(_i64(n), _i64(m), _d(np.ascontiguousarray(diag, dtype=np.float64)),
                               _i(np.ascontiguousarray(rows, dtype=np.int64)),
                               _i(np.ascontiguousarray(cols, dtype=np.int64)),
                               _d(np.ascontiguousarray(vals, dtype=np.float64)),
                               ctypes.c_int(int(dense_degree)), ctypes.c_double(omega),
                               ctypes.c_double(eta), _d(np.ascontiguousarray(x1, dtype=np.float64)),
                               _d(np.ascontiguousarray(y1, dtype=np.float64).reshape(-1)),
                               _d(x2), _d(y2))
    return x2, y2[:m]


def praq_canonical(n, diag, rows, cols, vals, dx, dy, omega=1.0, eta=0.999, dense_degree=False):
    f = lib().oracle_praq_canonical
    f.restype = ctypes.c_double
    return f(_i64(n), _i64(len(vals)), _d(np.ascontiguousarray(diag, dtype=np.float64)),
             _i(np.ascontiguousarray(rows, dtype=np.int64)),
             _i(np.ascontiguousarray(cols, dtype=np.int64)),
             _d(np.ascontiguousarray(vals, dtype=np.float64)), ctypes.c_int(int(dense_degree)),
             _d(np.ascontiguousarray(dx, dtype=np.float64)),
             _d(np.ascontiguousarray(dy, dtype=np.float64).reshape(-1)), ctypes.c_double(eta),
             ctypes.c_double(omega))


def simplex_project(v, target) -> np.ndarray:
    v = np.array(v, dtype=np.float64)
    lib().oracle_simplex_project(_i64(v.size), _d(v), ctypes.c_double(target))
    return v


def pria_solve(k, k0, h, params: OracleParams | None = None, warm=None) -> Solved:
    params = params or pria_params()
    k = np.ascontiguousarray(k, dtype=np.float64)
    h = np.ascontiguousarray(h, dtype=np.float64)
    n, t = h.shape[0], h.shape[1]
    x = np.zeros((n, t, t))
    y = np.zeros((n, t, t))
    z = np.zeros((n, t, t))
    w = [None, None, None]
    if warm is not None:
        w = [np.ascontiguousarray(a, dtype=np.float64) for a in warm]
    info = OracleInfo()
    st = lib().oracle_pria_solve(_i64(n), _i64(t), _d(k), ctypes.c_double(k0), _d(h),
                                 ctypes.byref(params), _d(w[0]), _d(w[1]), _d(w[2]),
                                 _d(x), _d(y), _d(z), ctypes.byref(info))
    if st < 0:
        raise ValueError(f"oracle_pria_solve rejected the instance ({st})")
    return Solved(x, y, z, info.iterations, info.restarts, info.lower, info.upper,
                  info.upper_iterate, info.relative_gap, info.status, info.scale, info.omega)


def pria_primal_objective(k, k0, h, x) -> float:
    h = np.ascontiguousarray(h, dtype=np.float64)
    n, t = h.shape[0], h.shape[1]
    return lib().oracle_pria_primal_objective(
        _i64(n), _i64(t), _d(np.ascontiguousarray(k, dtype=np.float64)), ctypes.c_double(k0),
        _d(h), _d(np.ascontiguousarray(x, dtype=np.float64)))


def pria_dual_objective(k, k0, h, y, z) -> float:
    h = np.ascontiguousarray(h, dtype=np.float64)
    n, t = h.shape[0], h.shape[1]
    return lib().oracle_pria_dual_objective(
        _i64(n), _i64(t), _d(np.ascontiguousarray(k, dtype=np.float64)), ctypes.c_double(k0),
        _d(h), _d(np.ascontiguousarray(y, dtype=np.float64)),
        _d(np.ascontiguousarray(z, dtype=np.float64)))
