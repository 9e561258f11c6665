"""Action / observation schemas (reference: coupledbubble_control/envs/common/specs.py:1-67).

A schema is an ordered list of components.  An observation component with
``stack`` s > 1 keeps the last s observations; its bounds are its lists
repeated s times (list repetition, as in the reference), while its values are
laid out bubble-major with the history innermost, so with per-bubble bounds
that differ, column j of such a component is scaled by bound j mod N.
``ActionMapingManager`` (the reference's spelling) maps action columns to
control_params rows and scales.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List

import torch

from .spaces import Box


@dataclass
class SpaceComponent:
    name: str
    idx: List[int]
    min: List[float]
    max: List[float]

    def get_bounds(self):
        return torch.as_tensor(self.min), torch.as_tensor(self.max)


@dataclass
class ObservationComponent(SpaceComponent):
    stack: int = 1

    def get_bounds(self):
        return torch.as_tensor(self.min * self.stack), torch.as_tensor(self.max * self.stack)


@dataclass
class ActionComponent(SpaceComponent):
    scale: float = 1.0


class SpaceSchema:
    components: List[SpaceComponent]

    def build_space(self, num_envs, device, **kwargs) -> Box:
        lows, highs = zip(*[c.get_bounds() for c in self.components])
        return Box(device=device, low=torch.cat(lows), high=torch.cat(highs), num_envs=num_envs, **kwargs)


class ActionMapingManager:
    """Column maps of the action vector (specs.py:45-67): ``_feat_col[name]`` the
    columns of a component, ``_dev_cols`` the control_params row of every column,
    ``_scales`` its scale."""

    def __init__(self, action_space_schema, control_layout) -> None:
        self.components = action_space_schema.components
        self.control_layout = control_layout
        self._feat_col = {}
        self._act_cols: List[int] = []
        self._dev_cols: List[int] = []
        self._scales: List[float] = []
        col = 0
        for comp in self.components:
            entry = control_layout.get(comp.name)
            if entry is None:
                raise KeyError(f"Missing control layout entry for action component '{comp.name}'")
            k = len(comp.idx)
            self._feat_col[comp.name] = range(col, col + k)
            self._act_cols.extend(range(col, col + k))
            self._dev_cols.extend(entry.base_offset + i for i in comp.idx)
            self._scales.extend([comp.scale] * k)
            col += k
