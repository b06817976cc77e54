"""Ragged batches of PRAQ and PRIA instances (include/reform_batch.h).

    b = Batch(use_float32=True)
    for problem in qubo_problems: b.add_praq(problem, praq_params)
    for problem in jrp_problems:  b.add_pria(problem, pria_params)
    b.upload()                      # host build + one transfer per instance
    seconds = b.solve("batched")    # all instances together
    x, y, info = b.praq_result(0)

"batched" advances every instance of a family with the same kernel launches
(block tables over the packed instances, per-instance restart state and
termination flags on the device, one check interval captured as a hipGraph
and replayed).  "sequential" runs the same resident instances one at a time
with the plain solver's tier choice: the baseline the batched mode is
measured against.  "workgroups" runs each instance's whole solve in one
workgroup of a persistent launch.  An instance's batched result does not
depend on the rest of its batch.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from . import _native, praq as _praq, pria as _pria

BATCHED, SEQUENTIAL, WORKGROUPS = 0, 1, 2
_MODES = {"batched": BATCHED, "sequential": SEQUENTIAL, "workgroups": WORKGROUPS}


class KernelStats(ctypes.Structure):
    """reform_kernel_stats of include/reform_batch.h."""
    _fields_ = [("name", ctypes.c_char * 64), ("launches", ctypes.c_int64),
                ("ms", ctypes.c_double), ("bytes", ctypes.c_double)]


def _declare(L):
    if getattr(L, "_batch_declared", False):
        return
    L.reform_batch_set_timing.restype = ctypes.c_int
    L.reform_batch_set_compaction.restype = ctypes.c_int
    L.reform_batch_set_compaction.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.reform_batch_set_timing.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.reform_batch_kernel_stats.restype = ctypes.c_int
    L.reform_batch_kernel_stats.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                            ctypes.POINTER(KernelStats)]
    L.reform_batch_create.restype = ctypes.c_void_p
    L.reform_batch_create.argtypes = [ctypes.c_int]
    L.reform_batch_destroy.argtypes = [ctypes.c_void_p]
    L.reform_batch_add_praq.restype = ctypes.c_int64
    L.reform_batch_add_praq.argtypes = [ctypes.c_void_p, ctypes.POINTER(_native.PraqInstance),
                                        ctypes.POINTER(_native.PraqParameters),
                                        ctypes.POINTER(_native.PraqOptions)]
    L.reform_batch_add_pria.restype = ctypes.c_int64
    L.reform_batch_add_pria.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                        _native._dp, ctypes.c_double, _native._dp,
                                        ctypes.POINTER(_native.PriaParams),
                                        ctypes.POINTER(_native.PriaOptions)]
    L.reform_batch_upload.restype = ctypes.c_int
    L.reform_batch_upload.argtypes = [ctypes.c_void_p]
    L.reform_batch_solve.restype = ctypes.c_int
    L.reform_batch_solve.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.reform_batch_set_shape.restype = ctypes.c_int
    L.reform_batch_set_shape.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    L.reform_batch_count.restype = ctypes.c_int64
    L.reform_batch_count.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.reform_batch_kernel_ms.restype = ctypes.c_int
    L.reform_batch_kernel_ms.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_double)]
    L.reform_batch_instance_seconds.restype = ctypes.c_double
    L.reform_batch_instance_seconds.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64]
    L.reform_batch_device_bytes.restype = ctypes.c_int64
    L.reform_batch_device_bytes.argtypes = [ctypes.c_void_p]
    L.reform_batch_praq_result.restype = ctypes.c_int
    L.reform_batch_praq_result.argtypes = [ctypes.c_void_p, ctypes.c_int64, _native._dp,
                                           _native._dp, ctypes.POINTER(_native.PraqInfo)]
    L.reform_batch_pria_result.restype = ctypes.c_int
    L.reform_batch_pria_result.argtypes = [ctypes.c_void_p, ctypes.c_int64, _native._dp,
                                           _native._dp, _native._dp,
                                           ctypes.POINTER(_native.PriaInfo)]
    L._batch_declared = True


def praq_native_params(params: _praq.Parameters):
    """PraqParameters / PraqOptions structs for a praq.Parameters."""
    L = _native.lib()
    p = L.praq_default_parameters()
    p.tolerance = params.tolerance
    p.max_iter = params.max_iter if params.max_iter > 0 else -1
    p.use_float32 = int(params.use_float32)
    p.omega = params.omega
    p.beta_sufficient = params.beta_sufficient
    p.beta_necessary = params.beta_necessary
    p.beta_artificial = params.beta_artificial
    p.theta_initial = params.theta
    p.theta_target = params.theta_target
    p.theta_decay = params.theta_decay
    p.op_steps = params.op_steps
    p.interval_offset = params.interval_offset
    p.interval_scale = params.interval_scale
    p.interval_power = params.interval_power
    o = L.praq_default_options()
    o.use_halpern = int(params.use_halpern)
    o.warm_steps = params.warm_steps
    o.tier = params.tier
    o.eta = params.eta
    o.time_limit = params.time_limit
    return p, o


class Batch:
    def __init__(self, use_float32: bool = True):
        self._L = _native.lib()
        _declare(self._L)
        self._h = self._L.reform_batch_create(int(use_float32))
        self.use_float32 = use_float32
        self._praq_shapes: list[tuple[int, int, bool]] = []
        self._pria_shapes: list[tuple[int, int]] = []

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.reform_batch_destroy(h)
            self._h = None

    def close(self):
        self.__del__()

    def add_praq(self, problem: _praq.QuboProblem, params: _praq.Parameters) -> int:
        p, o = praq_native_params(params)
        dense = problem.dense if params.dense_degree is None else params.dense_degree
        o.dense_degree = int(bool(dense))
        inst = _native.PraqInstance(problem.n(), problem.m(), _native.dptr(problem.diag),
                                    _native.iptr(problem.rows), _native.iptr(problem.cols),
                                    _native.dptr(problem.vals))
        q = self._L.reform_batch_add_praq(self._h, ctypes.byref(inst), ctypes.byref(p),
                                          ctypes.byref(o))
        if q < 0:
            raise ValueError(f"reform_batch_add_praq failed ({q})")
        self._praq_shapes.append((problem.n(), problem.m(), problem.dense))
        return int(q)

    def add_pria(self, problem: _pria.JrpProblem, params: _pria.Parameters) -> int:
        p = params.native()
        o = self._L.pria_default_options()
        o.tier = params.tier
        o.time_limit = params.time_limit
        k = np.ascontiguousarray(problem.order_costs, dtype=np.float64)
        h = np.ascontiguousarray(problem.holding_costs, dtype=np.float64)
        q = self._L.reform_batch_add_pria(self._h, problem.n(), problem.t(), _native.dptr(k),
                                          float(problem.joint_order_cost), _native.dptr(h),
                                          ctypes.byref(p), ctypes.byref(o))
        if q < 0:
            raise ValueError(f"reform_batch_add_pria failed ({q})")
        self._pria_shapes.append((problem.n(), problem.t()))
        return int(q)

    def set_shape(self, praq_block: int = 0, pria_block: int = 0) -> None:
        st = self._L.reform_batch_set_shape(self._h, praq_block, pria_block)
        if st < 0:
            raise ValueError(f"unsupported workgroup shape ({praq_block}, {pria_block})")

    def upload(self) -> None:
        st = self._L.reform_batch_upload(self._h)
        if st < 0:
            raise RuntimeError(f"reform_batch_upload failed ({st}): {_native.last_error()}")

    def solve(self, mode: str = "batched") -> float:
        """Solve every instance from its resident inputs; returns seconds."""
        t0 = time.perf_counter()
        st = self._L.reform_batch_solve(self._h, _MODES[mode])
        dt = time.perf_counter() - t0
        if st < 0:
            raise RuntimeError(f"reform_batch_solve failed ({st}): {_native.last_error()}")
        return dt

    def kernel_ms(self) -> tuple[float, float]:
        """(PRAQ, PRIA) launch durations of the last batched solve, by HIP
        events on the launch streams."""
        a, b = ctypes.c_double(), ctypes.c_double()
        self._L.reform_batch_kernel_ms(self._h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def set_compaction(self, on: bool = True) -> None:
        """Rebuild the batched mode's tables over the running instances as
        they finish (default on); results are the same either way."""
        self._L.reform_batch_set_compaction(self._h, int(on))

    def set_timing(self, on: bool = True) -> None:
        """Bracket every step launch of the batched mode with HIP events."""
        self._L.reform_batch_set_timing(self._h, int(on))

    def kernel_stats(self) -> list[dict]:
        """Per step kernel of the last batched solve (both families): name,
        launches, summed duration (ms) and algorithmic bytes."""
        out = []
        for family in (0, 1):
            arr = (KernelStats * 2)()
            st = self._L.reform_batch_kernel_stats(self._h, family, arr)
            if st < 0:
                raise RuntimeError(f"reform_batch_kernel_stats failed ({st}): "
                                   f"{_native.last_error()}")
            for k in arr:
                if k.launches:
                    out.append(dict(name=k.name.decode(), launches=int(k.launches),
                                    ms=float(k.ms), bytes=float(k.bytes)))
        return out

    def instance_seconds(self) -> tuple[np.ndarray, np.ndarray]:
        """(PRAQ, PRIA) per-instance host seconds of the last sequential
        solve, in add order."""
        np_, nr = self.counts()
        return (np.array([self._L.reform_batch_instance_seconds(self._h, 0, q)
                          for q in range(np_)]),
                np.array([self._L.reform_batch_instance_seconds(self._h, 1, q)
                          for q in range(nr)]))

    @property
    def device_bytes(self) -> int:
        return int(self._L.reform_batch_device_bytes(self._h))

    def counts(self) -> tuple[int, int]:
        return (int(self._L.reform_batch_count(self._h, 0)),
                int(self._L.reform_batch_count(self._h, 1)))

    def praq_result(self, q: int, with_solution: bool = True):
        n, m, _ = self._praq_shapes[q]
        x = np.zeros(n) if with_solution else None
        y = np.zeros((max(m, 1), 2)) if with_solution else None
        info = _native.PraqInfo()
        st = self._L.reform_batch_praq_result(self._h, q, _native.dptr(x), _native.dptr(y),
                                              ctypes.byref(info))
        if st < 0:
            raise RuntimeError(f"reform_batch_praq_result failed ({st}): {_native.last_error()}")
        return x, (y[:m] if y is not None else None), info

    def pria_result(self, q: int, with_solution: bool = True):
        n, t = self._pria_shapes[q]
        shape = (n, t, t)
        x = np.zeros(shape)
        y = np.zeros(shape) if with_solution else None
        z = np.zeros(shape) if with_solution else None
        info = _native.PriaInfo()
        st = self._L.reform_batch_pria_result(self._h, q, _native.dptr(x), _native.dptr(y),
                                              _native.dptr(z), ctypes.byref(info))
        if st < 0:
            raise RuntimeError(f"reform_batch_pria_result failed ({st}): {_native.last_error()}")
        return x, y, z, info
