"""Model base class: buffer layouts and the solve driver (reference: coupledbubble_control/models/model.py:1-285).

``Model.solve`` has the reference signature and calls the backend solver with
the reference keyword arguments (model.py:136-145), so callers such as the
vector environment's ``step_physics`` are unchanged.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Callable, Dict, Optional, Tuple

import numpy as np

from ..backends import BackendName, HIPOpts, KernelVariant
from .buffers import (ActualStateBuffer, BufferOps, BufferSet, DenseOutputBuffer, ParameterBuffer,
                      StatusBuffer)
from .specs import ExecutionSpec, LayoutEntry, ModelSpec, SolverSpec


@dataclass
class MaterialProperties:
    """Liquid properties in SI units.  Defaults are the reference's (model.py:16-24);
    unlike the reference they are dataclass fields, so they can be overridden."""
    PV: float = 0.0          # vapour pressure [Pa]
    RHO: float = 998.0       # liquid density [kg/m^3]
    ST: float = 0.0725       # surface tension [N/m]
    VIS: float = 0.001       # liquid viscosity [Pa s]
    CL: float = 1500.0       # liquid sound speed [m/s]
    P0: float = 1.013e5      # ambient pressure [Pa]
    PE: float = 1.4          # polytropic exponent [-]


def water_25c() -> MaterialProperties:
    """Water at 25 degC as used by the benchmark workloads."""
    return MaterialProperties(PV=3.17e3, RHO=998.0, ST=0.072, VIS=0.89e-3, CL=1497.0, P0=1.0e5, PE=1.4)


class ModelRuntime:
    def __init__(self, backend: BackendName, variant: KernelVariant, stream=None) -> None:
        self.backend = backend
        self.variant = variant
        self.stream = stream
        self.host_buffers: BufferSet
        self.device_buffers: Optional[BufferSet] = None
        self.solver: Optional[Callable] = None
        self.kernel: Any = None
        self.hip_opts: HIPOpts


class Model:
    _model_spec: ModelSpec
    _execution_spec: ExecutionSpec
    _solver_spec: SolverSpec
    _available_backends: Tuple[BackendName, ...]
    _kernel_variants: Tuple[KernelVariant, ...]
    runtime: ModelRuntime
    buffer_ops: BufferOps
    control_layout: Dict[str, LayoutEntry]
    state_layout: Dict[str, LayoutEntry]

    @property
    def time_shape(self):
        return (self._execution_spec.ns,)

    @property
    def state_shape(self):
        return (self._model_spec.sd, self._execution_spec.ns, self._execution_spec.tile)

    @property
    def unit_param_shape(self):
        return (self._model_spec.nup, self._execution_spec.ns, self._execution_spec.tile)

    @property
    def system_param_shape(self):
        return (self._model_spec.nsp, self._execution_spec.ns)

    @property
    def coupling_matrix_shape(self):
        return (self._model_spec.ncm, self._execution_spec.ups, self._execution_spec.ns, self._execution_spec.tile)

    @property
    def control_param_shape(self):
        return (self._model_spec.ncp, self._execution_spec.ns)

    @property
    def dense_state_shape(self):
        return (self._execution_spec.ndo, self._model_spec.sd, self._execution_spec.ns, self._execution_spec.tile)

    @property
    def dense_time_shape(self):
        return (self._execution_spec.ndo, self._execution_spec.ns)

    @property
    def dense_index_shape(self):
        return (self._execution_spec.ns,)

    @property
    def status_buffer_shape(self):
        return (self._execution_spec.ns,)

    def solve(self, max_steps: int = 100_000, kernel_steps: int = 2048, *, sync_to_host: bool = False,
              sync_from_host: bool = False, stream=None, benchmark: bool = False, debug: bool = False):
        """Integrate every system until its time_end or event (reference: model.py:108-161)."""
        self._ensure_buffers()
        if sync_from_host:
            self.buffer_ops.sync_to_device(self.runtime.host_buffers.state, self.runtime.device_buffers.state, stream)
            self.buffer_ops.sync_to_device(self.runtime.host_buffers.dense, self.runtime.device_buffers.dense, stream)
        if self.runtime.solver is None:
            self.runtime.solver = self._make_solver(self.runtime.backend, self.runtime.variant)
        out = self.runtime.solver(
            d_buffers=self.runtime.device_buffers,
            model_spec=self._model_spec,
            execution_spec=self._execution_spec,
            solver_spec=self._solver_spec,
            max_steps=max_steps,
            kernel_steps=kernel_steps,
            benchmark=benchmark,
            debug=debug,
        )
        if sync_to_host:
            self.buffer_ops.sync_to_host(self.runtime.device_buffers.state, self.runtime.host_buffers.state, stream)
            self.buffer_ops.sync_to_host(self.runtime.device_buffers.dense, self.runtime.host_buffers.dense, stream)
            self.buffer_ops.sync_to_host(self.runtime.device_buffers.status, self.runtime.host_buffers.status, stream)
        return out

    def build_parameters(self):
        raise NotImplementedError

    def build_control_parameters(self):
        raise NotImplementedError

    def build_default_initial_state(self):
        raise NotImplementedError

    def _initialize_host_buffers(self):
        """Host buffers with the reference defaults (model.py:180-224): h0 = 1e-6, t_end = 1,
        event = -1 (not initialised, avoids an immediate event termination)."""
        unit_params, system_params, coupling_matrices = self.build_parameters()
        control_params = self.build_control_parameters()
        params = ParameterBuffer(unit_params=unit_params, system_params=system_params,
                                 control_params=control_params, coupling_matrices=coupling_matrices)
        state = ActualStateBuffer(
            actual_time=np.zeros(self.time_shape, dtype=np.float64),
            time_end=np.ones(self.time_shape, dtype=np.float64),
            time_begin=np.zeros(self.time_shape, dtype=np.float64),
            time_step=np.full(self.time_shape, 1e-6, dtype=np.float64),
            actual_state=self.build_default_initial_state(),
        )
        dense = DenseOutputBuffer(
            dense_index=np.zeros(self.dense_index_shape, dtype=np.int32),
            dense_time=np.zeros(self.dense_time_shape, dtype=np.float64),
            dense_state=np.zeros(self.dense_state_shape, dtype=np.float64),
        )
        status = StatusBuffer(
            actual_event=np.full(self.status_buffer_shape, -1.0, dtype=np.float64),
            status_flags=np.zeros(self.status_buffer_shape, dtype=np.int32),
            total_steps=np.zeros(self.status_buffer_shape, dtype=np.int32),
            rejected_steps=np.zeros(self.status_buffer_shape, dtype=np.int32),
            convergece_success=np.zeros(self.status_buffer_shape, dtype=np.int32),
            convergence_failures=np.zeros(self.status_buffer_shape, dtype=np.int32),
        )
        self.runtime.host_buffers = BufferSet(state=state, params=params, dense=dense, status=status)

    def _ensure_buffers(self, stream=None):
        if self.runtime.device_buffers is not None:
            return
        hb = self.runtime.host_buffers
        self.runtime.device_buffers = BufferSet(
            state=self.buffer_ops.create_device_state(hb.state),
            params=self.buffer_ops.create_device_params(hb.params),
            dense=self.buffer_ops.create_device_dense_output(hb.dense),
            status=self.buffer_ops.create_device_status(hb.status),
        )

    def _check_shape(self, x: np.ndarray, expected: tuple, name: str) -> None:
        if x.shape != expected:
            raise ValueError(f"{name} shape mismatch: expected {expected}, got {x.shape}")

    def _shape_check_all(self, R0, FREQ, PA, PS):
        """Input shape checks of the reference (model.py:245-278)."""
        ns, ups, tile, nk = (self._execution_spec.ns, self._execution_spec.ups, self._execution_spec.tile,
                             self._model_spec.nk)
        R0 = np.asarray(R0, dtype=np.float64)
        FREQ = np.asarray(FREQ, dtype=np.float64)
        PA = np.asarray(PA, dtype=np.float64)
        PS = np.asarray(PS, dtype=np.float64)
        if R0.ndim != 2:
            raise ValueError(f"R0 must be 2D with shape ({ns}, {ups}) or ({ns}, {tile}), got {R0.shape}")
        for name, arr in (("FREQ", FREQ), ("PA", PA), ("PS", PS)):
            if arr.ndim != 2:
                raise ValueError(f"{name} must be 2D with shape ({nk}, {ns}), got {arr.shape}")
        if R0.shape not in ((ns, ups), (ns, tile)):
            raise ValueError(f"R0 shape mismatch: expected ({ns}, {ups}) or already padded ({ns}, {tile}), got {R0.shape}")
        for name, arr in (("FREQ", FREQ), ("PA", PA), ("PS", PS)):
            if arr.shape != (nk, ns):
                raise ValueError(f"{name} shape mismatch: expected {(nk, ns)}, got {arr.shape}")

    def _make_solver(self, b: BackendName, v: KernelVariant) -> Callable[..., Any]:
        raise NotImplementedError
