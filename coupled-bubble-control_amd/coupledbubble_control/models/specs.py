"""Model / execution / solver specifications (reference: coupledbubble_control/models/specs.py:1-78).

Same fields, defaults and derived values as the reference so that code which
builds or reads these specs is unchanged.  Two solver fields are added:
``col_threshold`` (the collision margin hard-coded as 0.1 in
backends/numba/sysmtem_definitions/ckm1d_def.py:534) and ``linsolve`` (the
coupling solve; "direct" = LU with partial pivoting).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Literal


@dataclass(frozen=True)
class ModelSpec:
    name: str           # Model name
    sd: int             # (unit) system dimension
    nup: int            # unit-specific parameters
    nsp: int            # shared (system) parameters
    ncm: int            # coupling matrices
    nk: int             # harmonic components
    ne: int             # events
    ncf: int            # coupling factors
    nct: int            # coupling terms
    ac: Literal["CONST", "SW_N", "SW_A"]    # acoustic field type
    ncp: int = field(init=False)            # 4 x NK (PA, FREQ, PS, WAVENUM)

    def __post_init__(self):
        object.__setattr__(self, "ncp", 4 * self.nk)


@dataclass(frozen=True)
class ExecutionSpec:
    ups: int                    # units (bubbles) per system
    ns: int = 1                 # number of systems
    spb: int = 1                # systems per block (reference launch shape; advisory here)
    ndo: int = 0                # dense outputs stored per system (0 = none)

    tile: int = field(init=False)           # ups padded to the next power of two
    total_threads: int = field(init=False)
    active_threads: int = field(init=False)
    block_size: int = field(init=False)
    grid_size: int = field(init=False)

    def __post_init__(self) -> None:
        if self.ups < 1:
            raise ValueError("ups must be > 1")
        if self.ns < 1:
            raise ValueError("ns must be >= 1")
        if self.spb < 1:
            raise ValueError("spb must be >= 1")
        object.__setattr__(self, "tile", 1 << (self.ups - 1).bit_length())
        object.__setattr__(self, "total_threads", self.ns * self.tile)
        object.__setattr__(self, "active_threads", self.ns * self.ups)
        object.__setattr__(self, "block_size", self.spb * self.tile)
        object.__setattr__(self, "grid_size", (self.ns + self.spb - 1) // self.spb)


@dataclass(frozen=True)
class SolverSpec:
    # -- ODE solver parameters --
    atol: float = 1e-9
    rtol: float = 1e-9
    min_step: float = 1e-16
    max_step: float = 1e-1
    growth_limit: float = 2.0
    shrink_limit: float = 0.5

    # -- coupling solve parameters --
    lin_atol: float = 1e-9
    lin_rtol: float = 1e-9
    lin_max_iter: int = 100
    lin_relaxation: float = 2.0 / 3.0
    etol: float = 1.0e-5

    # -- additions --
    col_threshold: float = 0.1
    linsolve: Literal["direct"] = "direct"
    # collision events: "bracket" (the reference GPU / Numba rule, |g| < etol
    # approached by shrinking rejected steps) or "exact" (the root of the gap on
    # the step's Hermite interpolant, as the reference's CPU path with solve_ivp
    # events; use with a small etol such as 1e-12)
    event_mode: Literal["bracket", "exact"] = "bracket"


@dataclass(frozen=True)
class LayoutEntry:
    group: str
    base_offset: int
