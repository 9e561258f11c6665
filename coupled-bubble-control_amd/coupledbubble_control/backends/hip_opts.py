"""Launch options of the HIP backend (the counterpart of backends/cuda_opts.py:5-73).

The reference's CUDAOpts carries NVRTC/NVCC compile flags.  Kernels here are
compiled ahead of time, so the options that remain are launch-shape choices.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Literal, Optional


@dataclass
class HIPOpts:
    mode: Literal["debug", "profile", "release"] = "release"
    block_size: Optional[int] = None   # threads per block; None = plan default
    verbose: bool = False              # print the plan (mapping, registers)
