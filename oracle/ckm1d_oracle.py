"""ctypes front end of the C oracle (oracle/ckm1d_oracle.c).  TEST INFRASTRUCTURE ONLY.

Runs the reference algorithm (Cash--Karp stepper of
backends/numba/rkck45_coupled_numba_stepper.py:116-706 with the CKM1D RHS of
ckm1d_def.py / ckm1d_def.cuh) on NumPy arrays in the reference buffer layout
(models/model.py:56-106), OpenMP over systems.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

from .build_oracle import ensure_built

FIELD = {"CONST": 0, "SW_N": 1, "SW_A": 2}
LIN = {"direct": 0, "richardson": 1}

_i32, _i64, _f64 = ctypes.c_int, ctypes.c_int64, ctypes.c_double
_pd = ctypes.POINTER(ctypes.c_double)
_pi = ctypes.POINTER(ctypes.c_int)
_pl = ctypes.POINTER(ctypes.c_int64)


class _Config(ctypes.Structure):
    _fields_ = [("ns", _i32), ("ups", _i32), ("tile", _i32), ("nk", _i32), ("field", _i32), ("ndo", _i32),
                ("atol", _f64), ("rtol", _f64), ("min_step", _f64), ("max_step", _f64), ("etol", _f64),
                ("linsolve", _i32), ("lin_atol", _f64), ("lin_rtol", _f64), ("lin_relax", _f64),
                ("lin_max_iter", _i32), ("col_threshold", _f64), ("event_exact", _i32)]


class _Tol(ctypes.Structure):
    _fields_ = [("atol", _f64), ("rtol", _f64), ("min_step", _f64), ("max_step", _f64), ("etol", _f64),
                ("event_exact", _i32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        h = ctypes.CDLL(str(ensure_built()))
        h.or_solve_reference_layout.restype = _i32
        h.or_solve_reference_layout.argtypes = [ctypes.POINTER(_Config), _i32, _i64,
                                                _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd,
                                                _pi, _pd, _pd, _pd, _pi, _pl, _pl]
        h.or_solve_periodic.restype = _i32
        h.or_solve_periodic.argtypes = [ctypes.POINTER(_Config), _i64, _f64, _i32, _i32,
                                        _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pd, _pi, _pl, _pl, _pd, _pd,
                                        _i32, _i32]
        h.or_rhs_reference_layout.restype = _i32
        h.or_rhs_reference_layout.argtypes = [ctypes.POINTER(_Config), _i32, _f64, _pd, _pd, _pd, _pd, _pd, _pd]
        h.or_event_reference_layout.restype = _f64
        h.or_event_reference_layout.argtypes = [ctypes.POINTER(_Config), _i32, _pd, _pd, _pd, _pd, _pd]
        h.or_tableau.restype = None
        h.or_tableau.argtypes = [_pd, _pd, _pd, _pd]
        h.or_fac_ok.restype = _f64
        h.or_fac_ok.argtypes = [_f64]
        h.or_fac_bad.restype = _f64
        h.or_fac_bad.argtypes = [_f64]
        h.or_set_threads.restype = _i32
        h.or_set_threads.argtypes = [_i32]
        h.or_linear_trace.restype = _i32
        h.or_linear_trace.argtypes = [_f64, _f64, _f64, _f64, ctypes.POINTER(_Tol), _i32, _pd, _pd, _pi]
        _lib = h
    return _lib


def _p(a: Optional[np.ndarray], ctype=_pd):
    if a is None:
        return None
    return a.ctypes.data_as(ctype)


@dataclass
class OracleSpec:
    ns: int
    ups: int
    nk: int
    field: str = "SW_N"
    ndo: int = 0
    atol: float = 1e-9
    rtol: float = 1e-9
    min_step: float = 1e-16
    max_step: float = 1e-1
    etol: float = 1e-5
    linsolve: str = "direct"
    lin_atol: float = 1e-9
    lin_rtol: float = 1e-9
    lin_relax: float = 1.0
    lin_max_iter: int = 100
    col_threshold: float = 0.1
    event_mode: str = "bracket"     # "bracket" (reference GPU / Numba rule) or "exact" (or_event_root)

    @property
    def tile(self) -> int:
        return 1 << (self.ups - 1).bit_length()

    def c(self) -> _Config:
        return _Config(self.ns, self.ups, self.tile, self.nk, FIELD[self.field], self.ndo, self.atol, self.rtol,
                       self.min_step, self.max_step, self.etol, LIN[self.linsolve], self.lin_atol, self.lin_rtol,
                       self.lin_relax, self.lin_max_iter, self.col_threshold, int(self.event_mode == "exact"))


def spec_from_model(model, linsolve: str = "direct", lin_relax: float = 1.0) -> OracleSpec:
    ms, es, ss = model._model_spec, model._execution_spec, model._solver_spec
    return OracleSpec(ns=es.ns, ups=es.ups, nk=ms.nk, field=ms.ac, ndo=es.ndo, atol=ss.atol, rtol=ss.rtol,
                      min_step=ss.min_step, max_step=ss.max_step, etol=ss.etol, linsolve=linsolve,
                      lin_atol=ss.lin_atol, lin_rtol=ss.lin_rtol, lin_relax=lin_relax,
                      lin_max_iter=ss.lin_max_iter, col_threshold=ss.col_threshold,
                      event_mode=getattr(ss, "event_mode", "bracket"))


def _c64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a


def set_threads(n: Optional[int]) -> int:
    """OpenMP threads of the oracle's loops over systems (n <= 0 or None: leave as is);
    returns the count in effect."""
    return int(lib().or_set_threads(int(n or 0)))


def solve(spec: OracleSpec, bufs: dict, kernel_steps: int = 2048, max_steps: int = 10 ** 9) -> dict:
    """Run the reference host loop on reference-layout NumPy buffers.  ``bufs`` holds
    actual_time, time_end, time_begin, time_step, actual_state, unit_params,
    system_params, control_params, coupling_matrices, actual_event, status_flags
    and, for ndo > 0, dense_index / dense_time / dense_state; the mutable ones are
    updated in place.  Returns the per-system attempt and rejection counts."""
    L = lib()
    cfg = spec.c()
    for k in ("actual_time", "time_end", "time_begin", "time_step", "actual_state", "unit_params",
              "system_params", "control_params", "coupling_matrices", "actual_event"):
        a = bufs[k]
        if a.dtype != np.float64 or not a.flags.c_contiguous:
            raise TypeError(f"{k} must be C-contiguous float64")
    if bufs["status_flags"].dtype != np.int32:
        raise TypeError("status_flags must be int32")
    total = np.zeros(spec.ns, dtype=np.int64)
    rej = np.zeros(spec.ns, dtype=np.int64)
    dense = spec.ndo > 0
    L.or_solve_reference_layout(
        ctypes.byref(cfg), int(kernel_steps), int(max_steps),
        _p(bufs["actual_time"]), _p(bufs["time_end"]), _p(bufs["time_begin"]), _p(bufs["time_step"]),
        _p(bufs["actual_state"]), _p(bufs["unit_params"]), _p(bufs["system_params"]),
        _p(bufs["control_params"]), _p(bufs["coupling_matrices"]),
        _p(bufs["dense_index"], _pi) if dense else None, _p(bufs["dense_time"]) if dense else None,
        _p(bufs["dense_state"]) if dense else None,
        _p(bufs["actual_event"]), _p(bufs["status_flags"], _pi), _p(total, _pl), _p(rej, _pl))
    return {"total_steps": total, "rejected_steps": rej}


def solve_periodic(spec: OracleSpec, bufs: dict, period: float, first_window: int, n_windows: int,
                   max_attempts: int = 10 ** 9, systems: Optional[slice] = None) -> dict:
    """Window-statistics integration to time_end (the parameter-study workload).  Returns
    counters plus max_r / mean_x shaped (n_windows, NS, TILE)."""
    L = lib()
    cfg = spec.c()
    total = np.zeros(spec.ns, dtype=np.int64)
    rej = np.zeros(spec.ns, dtype=np.int64)
    max_r = np.zeros((n_windows, spec.ns, spec.tile))
    mean_x = np.zeros((n_windows, spec.ns, spec.tile))
    s0, s1 = (0, spec.ns) if systems is None else (systems.start or 0, systems.stop or spec.ns)
    L.or_solve_periodic(
        ctypes.byref(cfg), int(max_attempts), float(period), int(first_window), int(n_windows),
        _p(bufs["actual_time"]), _p(bufs["time_end"]), _p(bufs["time_step"]), _p(bufs["actual_state"]),
        _p(bufs["unit_params"]), _p(bufs["system_params"]), _p(bufs["control_params"]),
        _p(bufs["coupling_matrices"]), _p(bufs["actual_event"]), _p(bufs["status_flags"], _pi),
        _p(total, _pl), _p(rej, _pl), _p(max_r), _p(mean_x), int(s0), int(s1))
    return {"total_steps": total, "rejected_steps": rej, "max_r": max_r, "mean_x": mean_x}


def solve_periodic_into(spec: OracleSpec, bufs: dict, period: float, first_window: int, n_windows: int,
                        max_r: np.ndarray, mean_x: np.ndarray, max_attempts: int) -> dict:
    """Like solve_periodic but accumulating into caller-owned window arrays, so that
    budget-limited calls resume open windows."""
    L = lib()
    cfg = spec.c()
    total = np.zeros(spec.ns, dtype=np.int64)
    rej = np.zeros(spec.ns, dtype=np.int64)
    assert max_r.shape == (n_windows, spec.ns, spec.tile) and max_r.flags.c_contiguous
    assert mean_x.shape == (n_windows, spec.ns, spec.tile) and mean_x.flags.c_contiguous
    L.or_solve_periodic(
        ctypes.byref(cfg), int(max_attempts), float(period), int(first_window), int(n_windows),
        _p(bufs["actual_time"]), _p(bufs["time_end"]), _p(bufs["time_step"]), _p(bufs["actual_state"]),
        _p(bufs["unit_params"]), _p(bufs["system_params"]), _p(bufs["control_params"]),
        _p(bufs["coupling_matrices"]), _p(bufs["actual_event"]), _p(bufs["status_flags"], _pi),
        _p(total, _pl), _p(rej, _pl), _p(max_r), _p(mean_x), 0, spec.ns)
    return {"total_steps": total, "rejected_steps": rej}


def rhs(spec: OracleSpec, bufs: dict, s: int, t: float, y: np.ndarray) -> np.ndarray:
    """Coupled RHS of system s at (t, y), y ordered [R(n), x(n), U(n), V(n)]."""
    L = lib()
    cfg = spec.c()
    y = _c64(y)
    dy = np.zeros_like(y)
    rc = L.or_rhs_reference_layout(ctypes.byref(cfg), int(s), float(t), _p(bufs["unit_params"]),
                                   _p(bufs["system_params"]), _p(bufs["control_params"]),
                                   _p(bufs["coupling_matrices"]), _p(y), _p(dy))
    if rc < 0:
        raise RuntimeError("oracle coupling solve failed")
    return dy


def event(spec: OracleSpec, bufs: dict, s: int, y: np.ndarray) -> float:
    L = lib()
    cfg = spec.c()
    y = _c64(y)
    return L.or_event_reference_layout(ctypes.byref(cfg), int(s), _p(bufs["unit_params"]),
                                       _p(bufs["system_params"]), _p(bufs["control_params"]),
                                       _p(bufs["coupling_matrices"]), _p(y))


def tableau():
    L = lib()
    a, b, c, cp = np.zeros(6), np.zeros(30), np.zeros(6), np.zeros(6)
    L.or_tableau(_p(a), _p(b), _p(c), _p(cp))
    return a, b.reshape(6, 5), c, cp


def fac_ok(r: float) -> float:
    return lib().or_fac_ok(float(r))


def fac_bad(r: float) -> float:
    return lib().or_fac_bad(float(r))


def linear_trace(lam: float, y0: float, t_end: float, h0: float, atol: float, rtol: float,
                 min_step: float = 1e-16, max_step: float = 0.1, max_attempts: int = 100000):
    """Attempt-by-attempt (t, h, accepted) of the stepper on y' = lam*y."""
    L = lib()
    tol = _Tol(atol, rtol, min_step, max_step, 1e-5)
    t = np.zeros(max_attempts)
    h = np.zeros(max_attempts)
    ok = np.zeros(max_attempts, dtype=np.int32)
    n = L.or_linear_trace(float(lam), float(y0), float(t_end), float(h0), ctypes.byref(tol), int(max_attempts),
                          _p(t), _p(h), _p(ok, _pi))
    return t[:n], h[:n], ok[:n].astype(bool)


def host_buffers_dict(model) -> dict:
    """Deep copy of a CKM1D model's host buffers as the dict ``solve`` expects."""
    hb = model.runtime.host_buffers
    d = {}
    for grp in (hb.state, hb.params, hb.dense, hb.status):
        for k in grp.__slots__:
            d[k] = np.array(getattr(grp, k), copy=True)
    return d
