"""Backend and kernel-variant selectors (reference: coupledbubble_control/backends/__init__.py:1-40).

The reference offers CuPy and Numba CUDA backends with "shared" and "warp"
kernel variants.  This package has one backend, HIP.  The reference variant
names are accepted and mean "let the plan choose" (AUTO), so existing
callers (e.g. the vector env's default ``KernelVariant.WARP``) keep working.
"""
from __future__ import annotations

from enum import Enum

from .hip_opts import HIPOpts


class BackendName(str, Enum):
    HIP = "hip"


class KernelVariant(str, Enum):
    AUTO = "auto"      # plan picks the mapping per bubble count (for N = 2 also by ensemble size)
    LANE = "lane"      # one system per lane
    GROUP = "group"    # TILE lanes per system, LDS-staged coupling solve
    # reference names (backends/__init__.py:8-10), both resolved like AUTO
    WARP = "warp"
    SHARED = "shared"


def get_current_device_name() -> str:
    """Sanitised name of the current HIP device (reference: backends/__init__.py:13-40)."""
    import re

    try:
        import torch

        if torch.cuda.is_available():
            name = torch.cuda.get_device_name()
            return re.sub(r"[^a-zA-Z0-9]", "", name).replace("AMD", "")
    except Exception:  # pragma: no cover - device query failure
        pass
    return "unknownGPU"


__all__ = ["BackendName", "KernelVariant", "HIPOpts", "get_current_device_name"]
