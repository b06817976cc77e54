"""ctypes binding of libreformulation_hip.so, the C ABI of include/*.h.

The library is built in-tree (``reformulation_amd/lib``) by
``__graft_entry__.build()``.  There is no fallback: every solve goes through
the HIP kernels, and a missing library or a missing GPU raises.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libreformulation_hip.so"

_i64 = ctypes.c_int64
_dbl = ctypes.c_double
_int = ctypes.c_int
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int64)


class PraqInstance(ctypes.Structure):
    _fields_ = [("n", _i64), ("m", _i64), ("diag", _dp), ("rows", _ip), ("cols", _ip),
                ("vals", _dp)]


class PraqParameters(ctypes.Structure):
    _fields_ = [("tolerance", _dbl), ("max_iter", _i64), ("use_float32", _int),
                ("omega", _dbl), ("beta_sufficient", _dbl), ("beta_necessary", _dbl),
                ("beta_artificial", _dbl), ("theta_initial", _dbl), ("theta_target", _dbl),
                ("theta_decay", _dbl), ("op_steps", _i64), ("interval_offset", _dbl),
                ("interval_scale", _dbl), ("interval_power", _dbl)]


class PraqResult(ctypes.Structure):
    _fields_ = [("lower", _dbl), ("upper", _dbl), ("iterations", _i64), ("status", _int)]


class PraqOptions(ctypes.Structure):
    _fields_ = [("use_halpern", _int), ("warm_steps", _i64), ("dense_degree", _int),
                ("tier", _int), ("eta", _dbl), ("time_limit", _dbl)]


class PraqInfo(ctypes.Structure):
    _fields_ = [("lower", _dbl), ("upper", _dbl), ("relative_gap", _dbl),
                ("iterations", _i64), ("restarts", _i64), ("checks", _i64),
                ("omega", _dbl), ("scale", _dbl), ("status", _int), ("tier", _int),
                ("switch_iteration", _i64), ("device_bytes", _i64), ("bracket", _dbl)]


class PriaParams(ctypes.Structure):
    _fields_ = [("omega", _dbl), ("eta", _dbl), ("beta_sufficient", _dbl),
                ("beta_artificial", _dbl), ("beta_necessary", _dbl), ("theta_initial", _dbl),
                ("theta_target", _dbl), ("theta_decay", _dbl), ("tolerance", _dbl),
                ("max_iter", _i64), ("use_halpern", _int), ("warm_steps", _int),
                ("op_steps", _int), ("verbose", _int)]


class PriaResultInfo(ctypes.Structure):
    _fields_ = [("iterations", _i64), ("primal_objective", _dbl), ("dual_objective", _dbl),
                ("relative_gap", _dbl)]


class PriaOptions(ctypes.Structure):
    _fields_ = [("tier", _int), ("time_limit", _dbl)]


class PriaInfo(ctypes.Structure):
    _fields_ = [("iterations", _i64), ("restarts", _i64), ("checks", _i64),
                ("lower", _dbl), ("upper", _dbl), ("upper_iterate", _dbl),
                ("relative_gap", _dbl), ("omega", _dbl), ("scale", _dbl), ("status", _int),
                ("tier", _int), ("switch_iteration", _i64), ("device_bytes", _i64),
                ("bracket", _dbl)]


#: Every symbol include/*.h declares; a test checks the library exports them.
EXPORTS = (
    "praq_default_parameters", "praq_check_interval", "praq_solve", "praq_default_options",
    "praq_solve_ex", "praq_solve_device",
    "pria_params_auto", "pria_solve_f64", "pria_solve_f32", "pria_solve_batch_f64",
    "pria_solve_batch_f32", "pria_objective_f64", "pria_default_options", "pria_solve_ex",
    "pria_solve_device", "pria_solve_device_noalloc",
    "reform_batch_create", "reform_batch_destroy", "reform_batch_add_praq",
    "reform_batch_add_pria", "reform_batch_upload", "reform_batch_solve",
    "reform_batch_set_shape", "reform_batch_count", "reform_batch_kernel_ms",
    "reform_batch_device_bytes", "reform_batch_praq_result", "reform_batch_pria_result",
    "reform_last_error", "reform_batch_set_timing", "reform_batch_kernel_stats",
    "reform_batch_set_compaction",
    "reform_batch_instance_seconds",
    "reform_tune_set", "reform_tune_get", "reform_tune_clear", "reform_praq_lanes_for",
    "reform_batch_praq_lanes_for",
    "reform_time_praq_steps",
    "reform_server_default_config", "reform_server_create", "reform_server_destroy",
    "reform_server_submit_praq", "reform_server_submit_pria", "reform_server_poll",
    "reform_server_wait_idle", "reform_server_info", "reform_server_praq_result",
    "reform_server_pria_result", "reform_server_release", "reform_server_stats",
    "reform_server_cancel", "reform_debug_fail_device_alloc", "reform_debug_device_allocs",
)

_lib = None


def last_error() -> str:
    """Text of the library's last failed HIP call ("" if none)."""
    return lib().reform_last_error().decode()


def available() -> bool:
    return LIB_PATH.exists()


def _torch_runtime_first() -> None:
    """torch's wheel ships its own HIP runtime under the system runtime's
    soname (libamdhip64.so.7), but its libraries name the file
    libamdhip64.so.  Loaded after this library (which names the soname),
    torch would map a second runtime into the process, and the two share no
    devices, streams or allocations: torch then finds no device.  Loaded
    first, torch's runtime is the one this library binds to, so tensors'
    data pointers and torch streams can be handed to the device-pointer entry
    points.  torch is imported here, when it is installed, for that order."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib() -> ctypes.CDLL:
    """Load the library, declaring argument and return types."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    _torch_runtime_first()
    L = ctypes.CDLL(str(LIB_PATH))
    L.reform_last_error.restype = ctypes.c_char_p
    L.reform_last_error.argtypes = []
    L.praq_default_parameters.restype = PraqParameters
    L.praq_default_options.restype = PraqOptions
    L.praq_check_interval.restype = _i64
    L.praq_check_interval.argtypes = [_i64, ctypes.POINTER(PraqParameters)]
    L.praq_solve.restype = _int
    L.praq_solve.argtypes = [ctypes.POINTER(PraqInstance), ctypes.POINTER(PraqParameters), _dp,
                             ctypes.POINTER(PraqResult)]
    L.praq_solve_ex.restype = _int
    L.praq_solve_ex.argtypes = [ctypes.POINTER(PraqInstance), ctypes.POINTER(PraqParameters),
                                ctypes.POINTER(PraqOptions), _dp, _dp, _dp, _dp,
                                ctypes.POINTER(PraqInfo)]
    # device pointers and the stream travel as plain addresses
    L.praq_solve_device.restype = _int
    L.praq_solve_device.argtypes = [ctypes.POINTER(PraqInstance), ctypes.POINTER(PraqParameters),
                                    ctypes.POINTER(PraqOptions), ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.POINTER(PraqInfo), ctypes.c_void_p]
    if hasattr(L, "pria_params_auto"):
        L.pria_params_auto.restype = PriaParams
        L.pria_params_auto.argtypes = [_i64, _dbl]
        L.pria_default_options.restype = PriaOptions
        L.pria_objective_f64.restype = _dbl
        L.pria_objective_f64.argtypes = [_i64, _i64, _dp, _dbl, _dp, _dp]
        for name, fp in (("pria_solve_f64", _dp), ("pria_solve_f32", _fp)):
            fn = getattr(L, name)
            fn.restype = _int
            real = _dbl if fp is _dp else ctypes.c_float
            fn.argtypes = [_i64, _i64, fp, real, fp, ctypes.POINTER(PriaParams), fp, fp, fp,
                           fp, fp, fp, ctypes.POINTER(PriaResultInfo)]
        for name, fp in (("pria_solve_batch_f64", _dp), ("pria_solve_batch_f32", _fp)):
            fn = getattr(L, name)
            fn.restype = _int
            fn.argtypes = [_i64, _i64, _i64, fp, fp, fp, ctypes.POINTER(PriaParams), fp, fp,
                           fp, ctypes.POINTER(PriaResultInfo)]
        L.pria_solve_ex.restype = _int
        L.pria_solve_ex.argtypes = [_i64, _i64, _dp, _dbl, _dp, ctypes.POINTER(PriaParams),
                                    ctypes.POINTER(PriaOptions), _int, _dp, _dp, _dp, _dp, _dp,
                                    _dp, ctypes.POINTER(PriaInfo)]
        for name in ("pria_solve_device", "pria_solve_device_noalloc"):
            fn = getattr(L, name)
            fn.restype = _int
            fn.argtypes = [_i64, _i64, ctypes.c_void_p, _dbl, ctypes.c_void_p,
                           ctypes.POINTER(PriaParams), ctypes.POINTER(PriaOptions), _int,
                           ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                           ctypes.POINTER(PriaInfo), ctypes.c_void_p]
    _lib = L
    from . import autotune        # the measured shape table, when one is present
    autotune.load(L=L)
    return L


def dptr(a):
    """double* of a C-contiguous float64 numpy array (or None)."""
    return None if a is None else a.ctypes.data_as(_dp)


def fptr(a):
    return None if a is None else a.ctypes.data_as(_fp)


def iptr(a):
    return None if a is None else a.ctypes.data_as(_ip)
