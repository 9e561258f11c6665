"""Serving handle of the HIP backend: one persistent kernel per model fed from a
host-mapped work ring (cbc_serve_* in include/cbc.h).

A submission is one environment step for a set of the model's systems: new
drive amplitudes (set_action, envs/bubble_1D_pos_control.py:731-749) and an
integration over tau in [0, t_end] from the resident state (step_physics and
advance_time, :751-753, 815-826).  Completion is read from host-mapped memory
(no device synchronisation): ``done(sys, seq)`` compares the published
sequence numbers, ``results(sys)`` copies the end states out of the
host-mapped slab.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np
import torch

from . import _abi
from .rkck45_coupled_hip_solver import _MAX_TRIG_ARG, _buffers_c, _check_shapes


@dataclass
class ServeOptions:
    capacity: int = 16384        # ring slots: items queued or running at once (power of two)
    waves: int = 0               # persistent grid in waves (0: lanes for min(capacity, ns) systems)
    refill: int = 1              # free lanes (lane groups) of a wave that trigger a claim
    poll_interval: int = 8       # attempts between ring polls of a wave with live systems
                                 # (up to 8x that after polls that claim nothing; 1 .. 2^20)
    max_attempts: int = 1 << 20  # per item; a system reaching it finishes ACTIVE
    idle_us: int = 200000        # idle waves exit after this long without work
    copy_state: bool = True      # copy end states to the host-mapped slab


def _host_array(addr: int, ctype, shape) -> np.ndarray:
    n = int(np.prod(shape))
    return np.ctypeslib.as_array((ctype * n).from_address(addr)).reshape(shape)


class ServeQueue:
    """Serving handle over a model's resident device buffers."""

    def __init__(self, model, opts: Optional[ServeOptions] = None, stream: Optional[torch.cuda.Stream] = None):
        """``stream=None`` lets the library create the serving stream with a hardware queue of
        its own (recommended: the serving kernel runs until it idles out)."""
        self._h = None
        self.model = model
        self.opts = opts or ServeOptions()
        model._ensure_buffers()
        if model.runtime.solver is None:
            model.runtime.solver = model._make_solver(model.runtime.backend, model.runtime.variant)
        self.kernel = model.runtime.kernel
        db = model.runtime.device_buffers
        _check_shapes(db, model._model_spec, model._execution_spec)
        self.ns = model._execution_spec.ns
        self.tile = model._execution_spec.tile
        self.ups = model._execution_spec.ups
        self.nk = model._model_spec.nk
        self.stream = stream
        cp = db.params.control_params
        sp = db.params.system_params
        self._wmax = float((cp[self.nk:2 * self.nk].abs() * sp[1:2].abs()).max().item()) * 2.0 * np.pi
        self._phmax = float(cp[2 * self.nk:3 * self.nk].abs().max().item())
        torch.cuda.current_stream().synchronize()
        o = self.opts
        spec = _abi.ServeSpecC(int(o.capacity), int(o.waves), int(o.refill), int(o.poll_interval),
                               int(o.max_attempts), int(o.idle_us), 1 if o.copy_state else 0)
        bufs = _buffers_c(db)
        h = ctypes.c_void_p()
        lib = _abi.lib()
        _abi.check(lib.cbc_serve_create(self.kernel.handle, ctypes.byref(bufs), ctypes.byref(spec),
                                        _abi.stream_handle(self.stream) if self.stream is not None else None,
                                        ctypes.byref(h)))
        self._h = h
        v = _abi.ServeViewC()
        _abi.check(lib.cbc_serve_get_view(h, ctypes.byref(v)))
        self.capacity = v.capacity
        self.waves = v.waves
        self._last_seq = np.zeros(self.ns, dtype=np.uint32)
        self.done_seq = _host_array(v.done_seq, ctypes.c_uint32, (self.ns,))
        self.out_misc = _host_array(v.out_misc, ctypes.c_double, (self.ns, 4))
        self.out_state = (_host_array(v.out_state, ctypes.c_double, (4, self.ns, self.tile))
                          if v.out_state else None)
        self.stage_state = _host_array(v.stage_state, ctypes.c_double, (4, self.ns, self.tile))
        self.stage_up = _host_array(v.stage_up, ctypes.c_double, (13, self.ns, self.tile))
        self.stage_cm = _host_array(v.stage_cm, ctypes.c_double, (3, self.ups, self.ns, self.tile))

    @property
    def handle(self) -> ctypes.c_void_p:
        if self._h is None:
            raise RuntimeError("serving handle is closed")
        return self._h

    def stage(self, sys: np.ndarray, state: np.ndarray, unit_params: np.ndarray, coupling: np.ndarray) -> None:
        """Write new-episode rows for systems ``sys`` (state (4, k, TILE), unit parameters
        (13, k, TILE), coupling (3, UPS, k, TILE)) into the staging mirrors; the next item of
        each system submitted with ``reset`` set takes them.  The systems must have no
        unfinished item."""
        sys = np.asarray(sys, dtype=np.int64)
        busy = ~self.done(sys, self._last_seq[sys])
        if busy.any():
            raise ValueError(f"system {int(sys[busy][0])} has an unfinished item")
        self.stage_state[:, sys] = state
        self.stage_up[:, sys] = unit_params
        self.stage_cm[:, :, sys] = coupling

    def submit(self, sys: np.ndarray, pa: Optional[np.ndarray], t_end: float, h0=0.0,
               reset: Optional[np.ndarray] = None, use_control: bool = False) -> Optional[np.ndarray]:
        """Publish one step for systems ``sys`` with amplitudes ``pa`` (NK, n) [Pa]; h0 > 0
        (scalar or per item) replaces the resident step size; ``reset`` (bool per item)
        starts new episodes from the staged rows; ``use_control`` takes the amplitudes from
        the device's control_params instead (``pa`` may then be None), where the
        environment's set_action has written them.  Returns the items' sequence numbers, or
        None when the ring has too few free slots (nothing published)."""
        sys = np.ascontiguousarray(sys, dtype=np.int32)
        n = sys.size
        if pa is None:
            if not use_control:
                raise ValueError("amplitudes are needed unless use_control is set")
            pa = np.zeros((self.nk, n))
        pa = np.ascontiguousarray(pa, dtype=np.float64).reshape(self.nk, n)
        if self._wmax * float(t_end) + self._phmax >= _MAX_TRIG_ARG:
            raise ValueError("t_end too long for the device sincos range")
        h = np.ascontiguousarray(np.broadcast_to(np.asarray(h0, dtype=np.float64), (n,)))
        opts = None
        if reset is not None or use_control:
            bits = np.zeros(n, dtype=np.int32)
            if reset is not None:
                bits |= np.where(np.asarray(reset, dtype=bool), _abi.CBC_SERVE_RESET, 0).astype(np.int32)
            if use_control:
                bits |= _abi.CBC_SERVE_USE_CONTROL
            opts = np.ascontiguousarray(bits)
        seq = np.empty(n, dtype=np.uint32)
        rc = _abi.lib().cbc_serve_submit(self.handle, sys.ctypes.data, pa.ctypes.data, int(n), float(t_end),
                                         h.ctypes.data, opts.ctypes.data if opts is not None else None,
                                         seq.ctypes.data)
        if rc == _abi.CBC_ERR_BUSY:
            return None
        _abi.check(rc)
        self._last_seq[sys] = seq
        return seq

    def poll(self) -> int:
        """Free finished ring slots and relaunch the kernel if it ended with work left;
        returns the number of items not yet known finished."""
        p = ctypes.c_int32(0)
        _abi.check(_abi.lib().cbc_serve_poll(self.handle, ctypes.byref(p)))
        return p.value

    def done(self, sys: np.ndarray, seq: np.ndarray) -> np.ndarray:
        """Per item: finished (its sequence number published)."""
        d = self.done_seq[np.asarray(sys)]
        return ((d.astype(np.int64) - np.asarray(seq, dtype=np.int64)) & 0xFFFFFFFF) < 0x80000000

    def results(self, sys: np.ndarray) -> Tuple[Optional[np.ndarray], np.ndarray]:
        """End states (4, n, UPS) (None without copy_state) and (n, 4) [tau, h, flag,
        attempts] of the last finished item of each system."""
        sys = np.asarray(sys)
        st = self.out_state[:, sys, :self.ups].copy() if self.out_state is not None else None
        return st, self.out_misc[sys].copy()

    def stats(self) -> Tuple[int, int, int]:
        launches = ctypes.c_int64(0)
        pub = ctypes.c_uint32(0)
        ret = ctypes.c_uint32(0)
        _abi.check(_abi.lib().cbc_serve_stats(self.handle, ctypes.byref(launches), ctypes.byref(pub),
                                              ctypes.byref(ret)))
        return launches.value, pub.value, ret.value

    def close(self) -> None:
        """Stop the kernel once its live items finish and free the handle."""
        if self._h is not None:
            h, self._h = self._h, None
            self.done_seq = self.out_misc = self.out_state = None
            self.stage_state = self.stage_up = self.stage_cm = None
            _abi.check(_abi.lib().cbc_serve_destroy(h))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
