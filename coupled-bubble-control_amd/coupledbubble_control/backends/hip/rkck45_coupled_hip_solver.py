"""Solver entry of the HIP backend (reference: backends/cupy_cuda/rkck45_coupled_cupy_solver.py:8-107,
backends/numba/rkck45_coupled_numba_solver.py:8-125).

``solver`` keeps the reference signature and semantics: up to
ceil(max_steps/kernel_steps) launches of kernel_steps step attempts each,
stopping as soon as no system is ACTIVE; with ``benchmark=True`` it returns
the list of per-launch wall seconds, else None.  The stop test reads one
device counter (systems left ACTIVE) instead of copying every status flag.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional

import numpy as np
import torch

from . import _abi

# trig arguments 2*pi*(f_k/f_ref)*tau + phase must stay below 2^20*pi/2 (device sincos range)
_MAX_TRIG_ARG = (2.0 ** 20) * (np.pi / 2.0)


def _buffers_c(d_buffers) -> _abi.BuffersC:
    st, pa, de, ss = d_buffers.state, d_buffers.params, d_buffers.dense, d_buffers.status
    for name, t in (("actual_state", st.actual_state), ("unit_params", pa.unit_params),
                    ("coupling_matrices", pa.coupling_matrices), ("control_params", pa.control_params)):
        if not isinstance(t, torch.Tensor) or not t.is_cuda:
            raise TypeError(f"{name} must be a torch tensor on the HIP device")
        if t.dtype != torch.float64 or not t.is_contiguous():
            raise TypeError(f"{name} must be contiguous float64")
    for name, t in (("status_flags", ss.status_flags), ("dense_index", de.dense_index)):
        if t.dtype != torch.int32:
            raise TypeError(f"{name} must be int32")
    p = _abi.ptr
    b = _abi.BuffersC()
    b.actual_time = p(st.actual_time)
    b.time_end = p(st.time_end)
    b.time_begin = p(st.time_begin)
    b.time_step = p(st.time_step)
    b.actual_state = p(st.actual_state)
    b.unit_params = p(pa.unit_params)
    b.system_params = p(pa.system_params)
    b.control_params = p(pa.control_params)
    b.coupling_matrices = p(pa.coupling_matrices)
    b.dense_index = p(de.dense_index) if de.dense_index.numel() else None
    b.dense_time = p(de.dense_time) if de.dense_time.numel() else None
    b.dense_state = p(de.dense_state) if de.dense_state.numel() else None
    b.actual_event = p(ss.actual_event)
    b.status_flags = p(ss.status_flags)
    b.total_steps = p(ss.total_steps) if ss.total_steps is not None and ss.total_steps.dtype == torch.int32 else None
    b.rejected_steps = (p(ss.rejected_steps)
                        if ss.rejected_steps is not None and ss.rejected_steps.dtype == torch.int32 else None)
    return b


def _check_shapes(d_buffers, model_spec, execution_spec):
    ns, tile, ups, nk = execution_spec.ns, execution_spec.tile, execution_spec.ups, model_spec.nk
    want = {
        "actual_state": (d_buffers.state.actual_state, (4, ns, tile)),
        "unit_params": (d_buffers.params.unit_params, (13, ns, tile)),
        "system_params": (d_buffers.params.system_params, (5, ns)),
        "control_params": (d_buffers.params.control_params, (4 * nk, ns)),
        "coupling_matrices": (d_buffers.params.coupling_matrices, (3, ups, ns, tile)),
        "actual_time": (d_buffers.state.actual_time, (ns,)),
        "time_end": (d_buffers.state.time_end, (ns,)),
        "time_step": (d_buffers.state.time_step, (ns,)),
        "actual_event": (d_buffers.status.actual_event, (ns,)),
        "status_flags": (d_buffers.status.status_flags, (ns,)),
    }
    if execution_spec.ndo > 0:
        want["dense_time"] = (d_buffers.dense.dense_time, (execution_spec.ndo, ns))
        want["dense_state"] = (d_buffers.dense.dense_state, (execution_spec.ndo, 4, ns, tile))
        want["dense_index"] = (d_buffers.dense.dense_index, (ns,))
    for name, (t, shape) in want.items():
        if tuple(t.shape) != shape:
            raise ValueError(f"{name} shape mismatch: expected {shape}, got {tuple(t.shape)}")


def _check_trig_range(d_buffers, model_spec):
    nk = model_spec.nk
    cp = d_buffers.params.control_params
    sp = d_buffers.params.system_params
    # one device-to-host read for the three maxima
    w, tm, ph = torch.stack([(cp[nk:2 * nk].abs() * sp[1:2].abs()).max(),
                             torch.maximum(d_buffers.state.time_end.abs(), d_buffers.state.actual_time.abs()).max(),
                             cp[2 * nk:3 * nk].abs().max()]).tolist()
    wmax, tmax, phmax = w * 2.0 * np.pi, tm, ph
    if wmax * tmax + phmax >= _MAX_TRIG_ARG:
        raise ValueError("time horizon too long for the device sincos range "
                         f"(|theta| up to {wmax * tmax + phmax:.3g} >= {_MAX_TRIG_ARG:.3g})")


def solver(kernel, d_buffers, model_spec, execution_spec, solver_spec, max_steps: int,
           kernel_steps: int = 2048, benchmark: bool = False, debug: bool = False,
           stream: Optional[torch.cuda.Stream] = None) -> Optional[List[float]]:
    lib = _abi.lib()
    _check_shapes(d_buffers, model_spec, execution_spec)
    _check_trig_range(d_buffers, model_spec)
    bufs = _buffers_c(d_buffers)
    max_launches = max(1, (int(max_steps) + int(kernel_steps) - 1) // int(kernel_steps))
    cap = min(max_launches, 1 << 16)
    times = (ctypes.c_double * cap)() if benchmark else None
    n = ctypes.c_int32(0)
    if debug:
        print("HIP plan", kernel.info())
    _abi.check(lib.cbc_solve(kernel.handle, ctypes.byref(bufs), int(kernel_steps), int(max_steps),
                             _abi.stream_handle(stream), times, cap if benchmark else 0, ctypes.byref(n)))
    if debug:
        print("launches", n.value)
    if benchmark:
        return [times[i] for i in range(min(n.value, cap))]
    return None


def launch(kernel, d_buffers, kernel_steps: int, stream=None) -> None:
    """One asynchronous launch of kernel_steps attempts (the reference kernel call itself)."""
    lib = _abi.lib()
    bufs = _buffers_c(d_buffers)
    _abi.check(lib.cbc_launch(kernel.handle, ctypes.byref(bufs), int(kernel_steps), _abi.stream_handle(stream)))


def solve_periodic(kernel, d_buffers, period: float, first_window: int, max_r: torch.Tensor,
                   mean_x: torch.Tensor, max_steps: int = 1_000_000, stream=None, resume: bool = False) -> None:
    """One launch integrating to time_end in windows of ``period`` and recording window
    statistics (asynchronous).  At most ``max_steps`` attempts per system: a system that
    runs out stays ACTIVE with its open window saved, and the next call (``resume=True``)
    continues it; such a continuation leaves systems that stopped on a failure (negative
    flag) untouched."""
    lib = _abi.lib()
    _check_shapes(d_buffers, kernel.model_spec, kernel.execution_spec)
    _check_trig_range(d_buffers, kernel.model_spec)
    es = kernel.execution_spec
    nw = int(max_r.shape[0])
    for name, t in (("max_r", max_r), ("mean_x", mean_x)):
        if tuple(t.shape) != (nw, es.ns, es.tile) or t.dtype != torch.float64 or not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous float64 of shape {(nw, es.ns, es.tile)}")
    bufs = _buffers_c(d_buffers)
    ps = _abi.PeriodSpecC(float(period), int(first_window), nw, _abi.ptr(max_r) if nw else None,
                          _abi.ptr(mean_x) if nw else None, int(bool(resume)))
    _abi.check(lib.cbc_solve_periodic(kernel.handle, ctypes.byref(bufs), ctypes.byref(ps), int(max_steps),
                                      _abi.stream_handle(stream)))


def horizon_order(d_buffers) -> torch.Tensor:
    """System indices by descending remaining horizon (time_end - actual_time), int32 on
    the device: the longest-processing-time-first order for the work queue."""
    rem = d_buffers.state.time_end - d_buffers.state.actual_time
    return torch.argsort(rem, descending=True, stable=True).to(torch.int32)


def active_horizon_order(d_buffers) -> torch.Tensor:
    """The systems still ACTIVE (flag 0), longest remaining horizon first (int32, device):
    the queue of a later launch round of one solve, which must leave systems that
    stopped -- on a failure included -- untouched.  May be empty."""
    idx = torch.nonzero(d_buffers.status.status_flags == 0).flatten()
    rem = (d_buffers.state.time_end - d_buffers.state.actual_time).index_select(0, idx)
    return idx.index_select(0, torch.argsort(rem, descending=True, stable=True)).to(torch.int32)


def _check_order(order: torch.Tensor, ns: int, validate: bool):
    if not isinstance(order, torch.Tensor) or order.dtype != torch.int32 or not order.is_cuda:
        raise TypeError("order must be an int32 tensor on the HIP device")
    if order.dim() != 1 or order.numel() < 1 or order.numel() > ns or not order.is_contiguous():
        raise ValueError(f"order must be a contiguous 1-D list of 1..{ns} system indices")
    if validate:
        if int(order.min().item()) < 0 or int(order.max().item()) >= ns:
            raise ValueError("order holds indices outside [0, ns)")
        if torch.unique(order).numel() != order.numel():
            raise ValueError("order holds duplicate system indices (two lanes would race on one system)")


def solve_queued(kernel, d_buffers, order: Optional[torch.Tensor] = None, waves: int = 0, refill: int = 1,
                 kernel_steps: int = 1 << 20, max_steps: int = 1 << 24, stream=None, validate: bool = True) -> int:
    """Persistent work-queue solve: lanes (or lane groups) take systems from ``order``
    (default: longest remaining horizon first) and refill as systems stop.  Relaunches
    while systems remain ACTIVE (at most ceil(max_steps/kernel_steps) launches).  Returns
    the launch count."""
    lib = _abi.lib()
    _check_shapes(d_buffers, kernel.model_spec, kernel.execution_spec)
    _check_trig_range(d_buffers, kernel.model_spec)
    ns = kernel.execution_spec.ns
    if order is None:
        order = horizon_order(d_buffers)
    _check_order(order, ns, validate)
    bufs = _buffers_c(d_buffers)
    n = ctypes.c_int32(0)
    _abi.check(lib.cbc_solve_queued(kernel.handle, ctypes.byref(bufs), _abi.ptr(order), int(order.numel()),
                                    int(waves), int(refill), int(kernel_steps), int(max_steps),
                                    _abi.stream_handle(stream), ctypes.byref(n)))
    return n.value


def launch_queued(kernel, d_buffers, order: torch.Tensor, waves: int = 0, refill: int = 1,
                  kernel_steps: int = 1 << 20, stream=None) -> None:
    """One asynchronous persistent launch (no host synchronisation; capturable).  The
    order tensor must stay alive until the launch has run."""
    lib = _abi.lib()
    bufs = _buffers_c(d_buffers)
    _abi.check(lib.cbc_launch_queued(kernel.handle, ctypes.byref(bufs), _abi.ptr(order), int(order.numel()),
                                     int(waves), int(refill), int(kernel_steps), _abi.stream_handle(stream)))


FUSED_MAX_BINS = 4   # bins of one fused launch (cbc_launch_fused)


def fused_supported(bubble_counts, nk: int, field: str) -> bool:
    """Whether one fused persistent kernel covers bins of these (distinct) bubble counts
    at `nk` harmonics and acoustic field `field` (ModelSpec.ac)."""
    ups = list(bubble_counts)
    if not ups or len(ups) > FUSED_MAX_BINS or len(set(ups)) != len(ups) or field not in _abi.FIELD_CODES:
        return False
    arr = (ctypes.c_int32 * len(ups))(*ups)
    return _abi.check(_abi.lib().cbc_fused_supported(arr, len(ups), int(nk), _abi.FIELD_CODES[field])) == 1


def launch_fused(kernels, d_buffers, orders, refills, waves: int = 0, kernel_steps: int = 1 << 20,
                 stream=None) -> None:
    """The queued launches of several bins (one plan per bubble count, group mapping) as
    ONE asynchronous persistent kernel: every wave drains the bins' queues in turn,
    largest bubble count first, and each bin's systems step exactly as in its own
    launch_queued.  Capturable; the order tensors must stay alive until it has run."""
    nb = len(kernels)
    if not (len(d_buffers) == len(orders) == len(refills) == nb):
        raise ValueError("one buffer set, order and refill per plan")
    lib = _abi.lib()
    bufs = [_buffers_c(b) for b in d_buffers]
    plans = (ctypes.c_void_p * nb)(*[k.handle.value for k in kernels])
    bptrs = (ctypes.POINTER(_abi.BuffersC) * nb)(*[ctypes.pointer(b) for b in bufs])
    optrs = (ctypes.c_void_p * nb)(*[_abi.ptr(o) for o in orders])
    lens = (ctypes.c_int32 * nb)(*[int(o.numel()) for o in orders])
    rf = (ctypes.c_int32 * nb)(*[int(r) for r in refills])
    _abi.check(lib.cbc_launch_fused(plans, bptrs, optrs, lens, rf, nb, int(waves), int(kernel_steps),
                                    _abi.stream_handle(stream)))


def active_count(kernel, stream=None) -> int:
    """Systems left ACTIVE by the plan's last launch (synchronises the stream)."""
    lib = _abi.lib()
    out = ctypes.c_int32(0)
    _abi.check(lib.cbc_plan_active_count(kernel.handle, _abi.stream_handle(stream), ctypes.byref(out)))
    return out.value
