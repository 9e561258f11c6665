"""Host and device buffer sets (reference: coupledbubble_control/models/buffers.py:1-200).

Same dataclasses and field names as the reference.  Host buffers are NumPy
arrays; device buffers are torch tensors on the current HIP device (the
reference used CuPy / Numba device arrays), so the environment's zero-copy
views (envs/common/interop.py:13-35) are the tensors themselves.
"""
from __future__ import annotations

from dataclasses import dataclass, fields
from typing import Any, Optional, Tuple

import numpy as np

from ..backends import BackendName


@dataclass(slots=True)
class Buffer:
    pass


@dataclass(slots=True)
class ActualStateBuffer(Buffer):
    actual_time: Any      # (NS,)
    time_end: Any         # (NS,)
    time_begin: Any       # (NS,)
    time_step: Any        # (NS,)
    actual_state: Any     # (SD, NS, TILE)


@dataclass(slots=True)
class ParameterBuffer(Buffer):
    unit_params: Any          # (NUP, NS, TILE)
    system_params: Any        # (NSP, NS)
    control_params: Any       # (NCP, NS)
    coupling_matrices: Any    # (NCM, UPS, NS, TILE)


@dataclass(slots=True)
class DenseOutputBuffer(Buffer):
    dense_index: Any      # (NS,) int32
    dense_time: Any       # (NDO, NS)
    dense_state: Any      # (NDO, SD, NS, TILE)


@dataclass(slots=True)
class StatusBuffer:
    actual_event: Any             # (NS,)
    status_flags: Any             # (NS,) int32
    total_steps: Any              # (NS,) int32, step attempts (filled by the HIP kernels)
    rejected_steps: Any           # (NS,) int32
    convergece_success: Any       # (NS,) int32 (kept for layout compatibility)
    convergence_failures: Any     # (NS,) int32


@dataclass(slots=True)
class BufferSet:
    state: ActualStateBuffer
    params: ParameterBuffer
    dense: DenseOutputBuffer
    status: StatusBuffer


class BufferOps:
    """Host <-> device buffer operations; backend None keeps NumPy arrays."""

    def __init__(self, backend: Optional[BackendName] = None, device: Optional[str] = None) -> None:
        self.backend = backend
        self.device = device or "cuda"

    def sync_to_device(self, src_buf, dst_buf, stream=None):
        for f in fields(src_buf):
            self._copy_h2d(getattr(src_buf, f.name), getattr(dst_buf, f.name), stream)

    def sync_to_host(self, src_buf, dst_buf, stream=None):
        for f in fields(src_buf):
            self._copy_d2h(getattr(src_buf, f.name), getattr(dst_buf, f.name), stream)

    def create_device_state(self, src_buf, stream=None):
        return ActualStateBuffer(**{f.name: self.from_host_ary(getattr(src_buf, f.name)) for f in fields(src_buf)})

    def create_device_params(self, src_buf, stream=None):
        return ParameterBuffer(**{f.name: self.from_host_ary(getattr(src_buf, f.name)) for f in fields(src_buf)})

    def create_device_dense_output(self, src_buf, stream=None):
        return DenseOutputBuffer(**{f.name: self.from_host_ary(getattr(src_buf, f.name)) for f in fields(src_buf)})

    def create_device_status(self, src_buf, stream=None):
        return StatusBuffer(**{f.name: self.from_host_ary(getattr(src_buf, f.name)) for f in fields(src_buf)})

    def from_host_ary(self, src, *, stream=None):
        host = np.ascontiguousarray(src)
        if self.backend == BackendName.HIP:
            import torch

            return torch.from_numpy(host.copy()).to(self.device)
        return host.copy()

    def _copy_h2d(self, src, dst, stream=None):
        if self.backend == BackendName.HIP:
            import torch

            dst.copy_(torch.from_numpy(np.ascontiguousarray(src)))
        else:
            np.copyto(dst, src)

    def _copy_d2h(self, src, dst, stream=None):
        if self.backend == BackendName.HIP:
            np.copyto(dst, src.detach().cpu().numpy())
        else:
            np.copyto(dst, src)

    @staticmethod
    def shape_of(x) -> Tuple[int, ...]:
        return tuple(x.shape)
