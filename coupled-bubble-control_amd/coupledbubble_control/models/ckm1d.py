"""Coupled Keller--Miksis bubble model with 1-D translation (reference: coupledbubble_control/models/ckm1d.py:1-255).

Same constructor, layouts and parameter construction as the reference; the
solver comes from the HIP backend.  Dimensionless variables: R/R0, x/lambda_ref,
U = dR/dt / (R0 f_ref), V = dx/dt / c; time tau = t f_ref (acoustic periods
of the reference frequency).
"""
from __future__ import annotations

from typing import Any, Callable, Literal, Optional

import numpy as np
from numpy.typing import NDArray

from ..backends import BackendName, HIPOpts, KernelVariant
from .buffers import BufferOps
from .model import MaterialProperties, Model, ModelRuntime
from .specs import ExecutionSpec, LayoutEntry, ModelSpec, SolverSpec


class CKM1D(Model):
    _available_backends = (BackendName.HIP,)
    _kernel_variants = (KernelVariant.AUTO, KernelVariant.LANE, KernelVariant.GROUP,
                        KernelVariant.WARP, KernelVariant.SHARED)

    def __init__(
            self,
            num_systems: int,
            systems_per_block: int,
            units_per_system: int,
            R0: NDArray[np.float64],   # (num_systems, units_per_system) [micron]
            PA: NDArray[np.float64],   # (num_components, num_systems)  [bar]
            FR: NDArray[np.float64],   # (num_components, num_systems)  [kHz]
            PS: NDArray[np.float64],   # (num_components, num_systems)  [-] (x pi rad)
            REF_FREQ: float = 20.0,    # kHz
            num_components: int = 2,
            acoustic_field: Literal["CONST", "SW_N", "SW_A"] = "CONST",
            mat_props: Optional[MaterialProperties] = None,
            backend: BackendName = BackendName.HIP,
            variant: KernelVariant = KernelVariant.AUTO,
            num_stored_points: int = 0,
            # accepted for call compatibility with the CUDA options of the reference
            compiler: str = "hipcc",
            max_registers: Optional[int] = None,
            fastmath: bool = False,
            cuda_mode: Literal["debug", "profile", "release"] = "release",
            solver_spec: Optional[SolverSpec] = None,
            block_size: Optional[int] = None,
            device: Optional[str] = None,
    ) -> None:
        if isinstance(backend, str) and not isinstance(backend, BackendName):
            backend = BackendName(backend)
        if backend not in self._available_backends:
            raise NotImplementedError(f"backend {backend!r} is not available; use BackendName.HIP")

        self._model_spec = ModelSpec(name="CKM1D", sd=4, nup=13, nsp=5, ncm=3, ne=1, ncf=4, nct=8,
                                     nk=num_components, ac=acoustic_field)
        self._execution_spec = ExecutionSpec(ns=num_systems, spb=systems_per_block, ups=units_per_system,
                                             ndo=num_stored_points)
        self._solver_spec = solver_spec if solver_spec is not None else SolverSpec()
        self._hip_opts = HIPOpts(mode=cuda_mode, block_size=block_size)
        self.runtime = ModelRuntime(backend, KernelVariant(variant))
        self.runtime.hip_opts = self._hip_opts
        self.buffer_ops = BufferOps(backend, device=device)

        nk = num_components
        self.control_layout = {
            "PA": LayoutEntry("control_params", 0),
            "FR": LayoutEntry("control_params", nk),
            "PS": LayoutEntry("control_params", 2 * nk),
            "WL": LayoutEntry("control_params", 3 * nk),
        }
        self.state_layout = {
            "R": LayoutEntry("actual_state", 0),
            "X": LayoutEntry("actual_state", 1),
            "U": LayoutEntry("actual_state", 2),
            "V": LayoutEntry("actual_state", 3),
        }
        self.mat_props = mat_props if mat_props is not None else MaterialProperties()
        self._shape_check_all(R0, FR, PA, PS)

        # SI units (ckm1d.py:113-127)
        self._REF_FREQ = REF_FREQ * 1000.0
        self._FREQ = np.ascontiguousarray(FR, dtype=np.float64) * 1000.0
        self._PA = np.ascontiguousarray(PA, dtype=np.float64) * 1.0e5
        self._PS = np.ascontiguousarray(PS, dtype=np.float64) * np.pi
        es = self._execution_spec
        R0 = np.asarray(R0, dtype=np.float64)
        if R0.shape == (es.ns, es.tile):
            self._R0 = np.ascontiguousarray(R0) * 1.0e-6
        else:
            self._R0 = np.zeros((es.ns, es.tile), dtype=np.float64)
            self._R0[:, :es.ups] = R0 * 1e-6
        self._initialize_host_buffers()

    # ------------------------------------------------------------------
    # parameter construction (ckm1d.py:133-212)
    # ------------------------------------------------------------------
    def build_parameters(self):
        """Unit parameters up[0..12], system parameters sp[0..4] and coupling
        matrices mx[0..2] of the reference (ckm1d.py:133-188)."""
        unit_params, coupling_matrices = self._unit_and_coupling(self._R0)
        system_params = np.zeros(self.system_param_shape, dtype=np.float64)
        TWO_PI = 2.0 * np.pi
        m = self.mat_props
        w_ref = TWO_PI * self._REF_FREQ
        l_ref = m.CL / self._REF_FREQ
        system_params[0] = 3.0 * m.PE
        system_params[1] = 1.0 / w_ref
        system_params[2] = l_ref / TWO_PI
        system_params[3] = m.CL
        system_params[4] = 1 / m.RHO / m.CL
        return unit_params, system_params, coupling_matrices

    def _unit_and_coupling(self, R0_si: np.ndarray):
        """Unit parameters (13, k, TILE) and coupling matrices (3, UPS, k, TILE) for k systems
        with equilibrium radii R0_si (k, TILE) in metres (ckm1d.py:146-188)."""
        k = R0_si.shape[0]
        es = self._execution_spec
        UPS = es.ups
        unit_params = np.zeros((self._model_spec.nup, k, es.tile), dtype=np.float64)
        coupling_matrices = np.zeros((self._model_spec.ncm, UPS, k, es.tile), dtype=np.float64)
        TWO_PI = 2.0 * np.pi
        m = self.mat_props
        PV, RHO, ST, VIS, CL, P0, PE = m.PV, m.RHO, m.ST, m.VIS, m.CL, m.P0, m.PE
        w_ref = TWO_PI * self._REF_FREQ
        l_ref = CL / self._REF_FREQ
        R0 = R0_si[:, :UPS]

        # Keller--Miksis
        unit_params[0, :, :UPS] = (2.0 * ST / R0 + P0 - PV) * (TWO_PI / R0 / w_ref) ** 2.0 / RHO
        unit_params[1, :, :UPS] = (1.0 - 3.0 * PE) * (2 * ST / R0 + P0 - PV) * (TWO_PI / R0 / w_ref) / CL / RHO
        unit_params[2, :, :UPS] = (P0 - PV) * (TWO_PI / R0 / w_ref) ** 2.0 / RHO
        unit_params[3, :, :UPS] = (2.0 * ST / R0 / RHO) * (TWO_PI / R0 / w_ref) ** 2.0
        unit_params[4, :, :UPS] = 4.0 * VIS / RHO / (R0 ** 2.0) * (TWO_PI / w_ref)
        unit_params[5, :, :UPS] = ((TWO_PI / R0 / w_ref) ** 2.0) / RHO
        unit_params[6, :, :UPS] = ((TWO_PI / w_ref) ** 2.0) / CL / RHO / R0
        unit_params[7, :, :UPS] = R0 * w_ref / TWO_PI / CL
        # translational motion
        unit_params[8, :, :UPS] = (0.5 * l_ref / R0) ** 2
        unit_params[9, :, :UPS] = TWO_PI / RHO / R0 / l_ref / (w_ref * R0) ** 2.0
        unit_params[10, :, :UPS] = 2 * TWO_PI / 3 * R0 ** 3
        unit_params[11, :, :UPS] = 6 * TWO_PI * VIS * R0
        unit_params[12, :, :UPS] = R0 / l_ref
        # coupling matrices: [c][row j][s][col i] = effect of bubble j on bubble i
        for row in range(UPS):
            Rj = R0[:, row:row + 1]
            coupling_matrices[0, row, :, :UPS] = Rj ** 3 / R0 ** 2 / l_ref
            coupling_matrices[1, row, :, :UPS] = 3.0 * (Rj / l_ref) ** 3 * np.ones_like(R0)
            coupling_matrices[2, row, :, :UPS] = (18 * VIS / RHO) * (TWO_PI / w_ref) * (Rj / l_ref) ** 3 / (R0 ** 2)
        idx = np.arange(UPS)
        coupling_matrices[:, idx, :, idx] = 0.0
        return unit_params, coupling_matrices

    def scene_rows(self, idx, R0: Optional[NDArray[np.float64]], X0: NDArray[np.float64]):
        """Rows of new episodes for systems ``idx``: new equilibrium radii R0 (k, UPS)
        [micron] (None keeps them; the host copy of R0 is updated), positions X0 (k, UPS)
        [wavelengths], R/R0 = 1, U = V = 0.  Returns (state (4, k, TILE), unit parameters
        (13, k, TILE), coupling (3, UPS, k, TILE))."""
        idx = np.asarray(idx, dtype=np.int64)
        k = idx.size
        es = self._execution_spec
        ups = es.ups
        X0 = np.asarray(X0, dtype=np.float64).reshape(k, ups)
        if R0 is not None:
            self._R0[idx, :ups] = np.asarray(R0, dtype=np.float64).reshape(k, ups) * 1e-6
        up, cm = self._unit_and_coupling(self._R0[idx])
        state = np.zeros((4, k, es.tile), dtype=np.float64)
        state[0, :, :ups] = 1.0
        state[1, :, :ups] = X0
        return state, up, cm

    def reset_systems(self, idx, R0: Optional[NDArray[np.float64]], X0: NDArray[np.float64],
                      h0: float = 1e-6, stream=None) -> None:
        """Start new episodes for systems ``idx`` (rows as in ``scene_rows``), with time 0,
        step size h0 and event value -1; the device rows are rewritten on ``stream``
        (reset_envs, envs/bubble_1D_pos_control.py:583-614, which keeps R0; the serving
        workload also draws new radii).  Asynchronous on ``stream``."""
        import torch

        idx = np.asarray(idx, dtype=np.int64)
        state, up, cm = self.scene_rows(idx, R0, X0)
        self._ensure_buffers()
        db = self.runtime.device_buffers
        dev = db.state.actual_state.device
        st = stream if stream is not None else torch.cuda.current_stream(dev)
        with torch.cuda.stream(st):
            ti = torch.from_numpy(idx).to(dev)
            db.params.unit_params.index_copy_(1, ti, torch.from_numpy(up).to(dev))
            db.params.coupling_matrices.index_copy_(2, ti, torch.from_numpy(cm).to(dev))
            db.state.actual_state.index_copy_(1, ti, torch.from_numpy(state).to(dev))
            db.state.actual_time.index_fill_(0, ti, 0.0)
            db.state.time_step.index_fill_(0, ti, float(h0))
            db.status.actual_event.index_fill_(0, ti, -1.0)

    def build_control_parameters(self):
        """[PA (Pa), omega (rad/s), phase (rad), wave number (rad/m)] x NK (ckm1d.py:190-202)."""
        control_params = np.zeros(self.control_param_shape, dtype=np.float64)
        TWO_PI = 2.0 * np.pi
        CL = self.mat_props.CL
        NC = self._model_spec.nk
        control_params[0:NC, :] = self._PA
        control_params[NC:2 * NC, :] = self._FREQ * TWO_PI
        control_params[2 * NC:3 * NC, :] = self._PS
        control_params[3 * NC:4 * NC, :] = TWO_PI * self._FREQ / CL
        return control_params

    def build_default_initial_state(self):
        """R/R0 = 1, x = U = V = 0 for every real unit (ckm1d.py:204-212)."""
        state = np.zeros(self.state_shape, dtype=np.float64)
        state[0, :, :self._execution_spec.ups] = 1.0
        return state

    # ------------------------------------------------------------------
    def _make_solver(self, b: BackendName, v: KernelVariant = KernelVariant.AUTO) -> Callable[..., Any]:
        """Backend solver bound to this model's plan (ckm1d.py:214-253)."""
        from functools import partial

        if b != BackendName.HIP:
            raise NotImplementedError(f"backend {b!r}")
        from ..backends.hip.kernel_loader import load_coupled_kernel
        from ..backends.hip.rkck45_coupled_hip_solver import solver

        kernel = load_coupled_kernel(self._model_spec, self._execution_spec, self._solver_spec,
                                     self._hip_opts, variant=v)
        self.runtime.kernel = kernel
        return partial(solver, kernel=kernel)

    def solve_periodic(self, period: float = 1.0, first_window: int = 0, n_windows: Optional[int] = None,
                       max_steps: int = 1_000_000, stream=None, max_launches: int = 1000):
        """Integrate every system to its time_end in windows of ``period`` and return
        (max_r, mean_x) shaped (n_windows, NS, TILE): the largest R/R0 over the
        accepted states of each window and the trapezoidal time-mean of x/lambda.
        Windows with index < first_window are skipped.  Launches of at most
        ``max_steps`` attempts per system repeat while any system is ACTIVE (at most
        ``max_launches``).  Extension of the reference, whose parameter study derives
        such outputs from dense output."""
        import torch

        from ..backends.hip.rkck45_coupled_hip_solver import solve_periodic

        self._ensure_buffers()
        if self.runtime.solver is None:
            self.runtime.solver = self._make_solver(self.runtime.backend, self.runtime.variant)
        db = self.runtime.device_buffers
        if n_windows is None:
            t_end = float(db.state.time_end.max().item())
            n_windows = max(0, int(np.ceil(t_end / period - 1e-9)) - first_window)
        es = self._execution_spec
        dev = db.state.actual_state.device
        max_r = torch.zeros((n_windows, es.ns, es.tile), dtype=torch.float64, device=dev)
        mean_x = torch.zeros((n_windows, es.ns, es.tile), dtype=torch.float64, device=dev)
        for k in range(max_launches):
            solve_periodic(self.runtime.kernel, db, period, first_window, max_r, mean_x, max_steps, stream,
                           resume=k > 0)
            if not bool((db.status.status_flags == 0).any().item()):
                break
        return max_r, mean_x
