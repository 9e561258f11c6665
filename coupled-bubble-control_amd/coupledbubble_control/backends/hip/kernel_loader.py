"""Plan creation for the HIP backend (replaces backends/cupy_cuda/kernel_loader.py:9-99).

The reference JIT-compiles one CUDA kernel per spec with every constant as a
``-D`` macro.  Here every kernel is compiled ahead of time into libcbc.so and
``cbc_plan_create`` selects the instantiation (bubble count, harmonics, field
type) and the mapping; the spec values travel as launch arguments.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional

from .. import HIPOpts, KernelVariant
from . import _abi

_MAPPINGS = {
    KernelVariant.AUTO: _abi.CBC_MAP_AUTO,
    KernelVariant.WARP: _abi.CBC_MAP_AUTO,
    KernelVariant.SHARED: _abi.CBC_MAP_AUTO,
    KernelVariant.LANE: _abi.CBC_MAP_LANE,
    KernelVariant.GROUP: _abi.CBC_MAP_GROUP,
}


@dataclass
class PlanInfo:
    mapping: str
    block_size: int
    grid_size: int
    systems_per_block: int
    lds_bytes: int
    num_regs: int
    local_bytes: int


class HipKernel:
    """Owns one cbc_plan; destroyed with the object."""

    def __init__(self, model_spec, execution_spec, solver_spec, opts: Optional[HIPOpts] = None,
                 variant: KernelVariant = KernelVariant.AUTO):
        lib = _abi.lib()
        opts = opts or HIPOpts()
        if model_spec.ac not in _abi.FIELD_CODES:
            raise ValueError(f"unknown acoustic field {model_spec.ac!r}")
        if solver_spec.linsolve != "direct":
            raise ValueError(f"unsupported coupling solve {solver_spec.linsolve!r}")
        self.model_spec = model_spec
        self.execution_spec = execution_spec
        self.solver_spec = solver_spec
        m = _abi.ModelSpecC(model_spec.sd, model_spec.nup, model_spec.nsp, model_spec.ncm, model_spec.nk,
                            model_spec.ne, model_spec.ncf, model_spec.nct, _abi.FIELD_CODES[model_spec.ac])
        e = _abi.ExecSpecC(execution_spec.ns, execution_spec.spb, execution_spec.ups, execution_spec.ndo,
                           execution_spec.tile, _MAPPINGS[KernelVariant(variant)], int(opts.block_size or 0))
        s = _abi.SolverSpecC(solver_spec.atol, solver_spec.rtol, solver_spec.min_step, solver_spec.max_step,
                             solver_spec.etol, solver_spec.lin_atol, solver_spec.lin_rtol,
                             solver_spec.lin_relaxation, solver_spec.lin_max_iter, _abi.CBC_LIN_DIRECT,
                             solver_spec.col_threshold, _abi.EVENT_MODES[solver_spec.event_mode])
        handle = ctypes.c_void_p()
        _abi.check(lib.cbc_plan_create(ctypes.byref(m), ctypes.byref(e), ctypes.byref(s), ctypes.byref(handle)))
        self._handle = handle
        self._lib = lib
        if opts.verbose:
            print(self.info())

    @property
    def handle(self) -> ctypes.c_void_p:
        if self._handle is None:
            raise RuntimeError("plan already destroyed")
        return self._handle

    def info(self) -> PlanInfo:
        pi = _abi.PlanInfoC()
        _abi.check(self._lib.cbc_plan_query(self.handle, ctypes.byref(pi)))
        names = {_abi.CBC_MAP_LANE: "lane", _abi.CBC_MAP_GROUP: "group"}
        return PlanInfo(names.get(pi.mapping, str(pi.mapping)), pi.block_size, pi.grid_size, pi.systems_per_block,
                        pi.lds_bytes, pi.num_regs, pi.local_bytes)

    def close(self):
        if getattr(self, "_handle", None) is not None:
            self._lib.cbc_plan_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def load_coupled_kernel(model_spec, execution_spec, solver_spec, opts: Optional[HIPOpts] = None,
                        variant: KernelVariant = KernelVariant.AUTO) -> HipKernel:
    """Select and size the kernel for the given specs (reference signature minus the CUDA source names)."""
    return HipKernel(model_spec, execution_spec, solver_spec, opts, variant)
