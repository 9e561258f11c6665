"""Synthetic ensembles of the benchmark and parity workloads.

Every generator is seeded (numpy default_rng) and returns a :class:`Workload`
holding physical inputs in the reference's units (micron, bar, kHz, phase in
units of pi; models/ckm1d.py:38-42), initial positions in wavelengths of the
reference frequency, and per-system horizons in acoustic periods.

* ``plain_cpu``      256 systems of N=2, dual-frequency standing wave (25/50 kHz),
                     1 RL step of 5 periods, rtol=atol=1e-9.
* ``param_study``    65 536 systems of N=2 on a 256 x 256 (PA1, PA2) grid over
                     [0, 2] bar, R0 = (44, 60) um, 200 periods with the first
                     50 discarded from the per-period statistics, rtol=atol=1e-10.
* ``mixed_uneven``   32 768 systems with N in {2, 3, 4, 8} at 8:4:2:1, uneven horizons.
* ``serving_bins``   16 384 resident systems, N in {2, 4} at 3:1, c0-style scenes,
                     5-period steps (the serving loop's envs).
* ``large_coupling`` 4 096 systems of N=16 mixed 1:1 with N=2 systems.
Water at 25 degC throughout.
"""
from __future__ import annotations

from dataclasses import dataclass, field as dc_field
from typing import List, Optional

import numpy as np

from .models.model import MaterialProperties, water_25c


@dataclass
class Workload:
    name: str
    n: int                      # bubbles per system
    R0: np.ndarray              # (NS, n) micron
    X0: np.ndarray              # (NS, n) x / lambda_ref
    PA: np.ndarray              # (NK, NS) bar
    FR: np.ndarray              # (NK, NS) kHz
    PS: np.ndarray              # (NK, NS) units of pi
    t_end: np.ndarray           # (NS,) periods of the reference frequency
    ref_freq: float = 25.0      # kHz
    field: str = "SW_N"
    rtol: float = 1e-9
    atol: float = 1e-9
    mat: MaterialProperties = dc_field(default_factory=water_25c)
    period: float = 0.0         # statistics window length (0 = none)
    first_window: int = 0
    n_windows: int = 0

    @property
    def ns(self) -> int:
        return int(self.R0.shape[0])

    @property
    def nk(self) -> int:
        return int(self.PA.shape[0])

    def subset(self, idx) -> "Workload":
        idx = np.asarray(idx)
        return Workload(self.name, self.n, self.R0[idx], self.X0[idx], self.PA[:, idx], self.FR[:, idx],
                        self.PS[:, idx], self.t_end[idx], self.ref_freq, self.field, self.rtol, self.atol,
                        self.mat, self.period, self.first_window, self.n_windows)


def wavelength(mat: MaterialProperties, ref_freq_khz: float) -> float:
    return mat.CL / (ref_freq_khz * 1e3)


def sample_positions(rng, n_systems: int, n: int, lo: float, hi: float, min_gap: np.ndarray) -> np.ndarray:
    """Sorted uniform positions on [lo, hi] (wavelengths), rejection-sampled per system
    until every neighbour gap is at least min_gap[s]."""
    out = np.empty((n_systems, n))
    todo = np.arange(n_systems)
    while todo.size:
        x = np.sort(rng.uniform(lo, hi, size=(todo.size, n)), axis=1)
        ok = np.all(np.diff(x, axis=1) >= min_gap[todo, None], axis=1) if n > 1 else np.ones(todo.size, bool)
        out[todo[ok]] = x[ok]
        todo = todo[~ok]
    return out


def plain_cpu(seed: int = 0, ns: int = 256) -> Workload:
    rng = np.random.default_rng(seed)
    mat = water_25c()
    lam = wavelength(mat, 25.0)
    R0 = rng.uniform(20.0, 80.0, size=(ns, 2))
    gap = 3.0 * R0.max(axis=1) * 1e-6 / lam
    X0 = sample_positions(rng, ns, 2, 0.0, 0.25, gap)
    PA = rng.uniform(0.0, 1.5, size=(2, ns))
    FR = np.tile(np.array([[25.0], [50.0]]), (1, ns))
    PS = np.zeros((2, ns))
    return Workload("plain_cpu_256xN2", 2, R0, X0, PA, FR, PS, np.full(ns, 5.0), mat=mat)


def param_study(side: int = 256, periods: float = 200.0, discard: int = 50, seed: int = 1) -> Workload:
    mat = water_25c()
    lam = wavelength(mat, 25.0)
    ns = side * side
    pa = np.linspace(0.0, 2.0, side)
    P1, P2 = np.meshgrid(pa, pa, indexing="ij")
    PA = np.stack([P1.ravel(), P2.ravel()])
    R0 = np.tile(np.array([44.0, 60.0]), (ns, 1))
    rng = np.random.default_rng(seed)
    gap = np.array([3.0 * 60.0 * 1e-6 / lam])
    x_pair = sample_positions(rng, 1, 2, 0.0, 0.25, gap)[0]
    X0 = np.tile(x_pair, (ns, 1))
    FR = np.tile(np.array([[25.0], [50.0]]), (1, ns))
    PS = np.zeros((2, ns))
    n_windows = int(round(periods)) - discard
    return Workload(f"param_study_{ns}xN2", 2, R0, X0, PA, FR, PS, np.full(ns, float(periods)),
                    rtol=1e-10, atol=1e-10, mat=mat, period=1.0, first_window=discard, n_windows=n_windows)


def mixed_uneven(seed: int = 2, total: int = 32768) -> List[Workload]:
    """Per-N sub-ensembles (N = 2, 3, 4, 8 at 8:4:2:1) with log-uniform R0 in [10, 100] um,
    sorted positions on [0, lambda/2] with gaps >= 3 max(R0), horizons in [1, 50] periods."""
    rng = np.random.default_rng(seed)
    mat = water_25c()
    lam = wavelength(mat, 25.0)
    weights = {2: 8, 3: 4, 4: 2, 8: 1}
    wsum = sum(weights.values())
    counts = {n: total * w // wsum for n, w in weights.items()}
    counts[2] += total - sum(counts.values())
    out = []
    for n, ns in counts.items():
        R0 = np.exp(rng.uniform(np.log(10.0), np.log(100.0), size=(ns, n)))
        gap = 3.0 * R0.max(axis=1) * 1e-6 / lam
        X0 = sample_positions(rng, ns, n, 0.0, 0.5, gap)
        PA = rng.uniform(0.0, 1.5, size=(2, ns))
        FR = np.tile(np.array([[25.0], [50.0]]), (1, ns))
        t_end = rng.integers(1, 51, size=ns).astype(np.float64)
        out.append(Workload(f"mixed_N{n}", n, R0, X0, PA, FR, np.zeros((2, ns)), t_end, mat=mat))
    return out


def serving_bins(seed: int = 3, sizes: Optional[dict] = None, periods: float = 5.0) -> dict:
    """Resident envs of the serving workload, one Workload per bubble count (default
    12 288 systems of N=2 and 4 096 of N=4): R0 ~ U[20, 80] um, positions sorted
    U[0, 0.25] wavelengths with gaps >= 3 max(R0), drive 25/50 kHz with
    PA ~ U[0, 1.5] bar, K = ``periods`` per step, rtol = atol = 1e-9."""
    rng = np.random.default_rng(seed)
    mat = water_25c()
    lam = wavelength(mat, 25.0)
    sizes = sizes or {2: 12288, 4: 4096}
    out = {}
    for n, ns in sizes.items():
        R0 = rng.uniform(20.0, 80.0, size=(ns, n))
        gap = 3.0 * R0.max(axis=1) * 1e-6 / lam
        X0 = sample_positions(rng, ns, n, 0.0, 0.25, gap)
        PA = rng.uniform(0.0, 1.5, size=(2, ns))
        FR = np.tile(np.array([[25.0], [50.0]]), (1, ns))
        out[n] = Workload(f"serving_N{n}", n, R0, X0, PA, FR, np.zeros((2, ns)), np.full(ns, periods), mat=mat)
    return out


def large_coupling(seed: int = 4, ns16: int = 4096, ns2: int = 4096) -> List[Workload]:
    rng = np.random.default_rng(seed)
    mat = water_25c()
    lam = wavelength(mat, 25.0)
    n = 16
    R0 = np.exp(rng.uniform(np.log(5.0), np.log(50.0), size=(ns16, n)))
    rmax = R0.max(axis=1) * 1e-6 / lam
    j = (1.0 / n - 3.0 * rmax) / 2.0
    base = np.arange(n) / n + 0.5 / n
    X0 = base[None, :] + rng.uniform(-1.0, 1.0, size=(ns16, n)) * j[:, None]
    PA = rng.uniform(0.5, 2.0, size=(2, ns16))
    FR = np.tile(np.array([[25.0], [37.5]]), (1, ns16))
    w16 = Workload("large_N16", 16, R0, X0, PA, FR, np.zeros((2, ns16)), np.full(ns16, 20.0), mat=mat)
    w2 = plain_cpu(seed=seed + 100, ns=ns2)
    w2.name = "large_mix_N2"
    w2.FR = np.tile(np.array([[25.0], [37.5]]), (1, ns2))
    w2.t_end = np.full(ns2, 20.0)
    return [w16, w2]


# The benchmark harness's scenes (scripts/benchmarks/kernel_prof.py:65-94): R0 60 um, PA of
# both harmonics ~ U[0, 1] bar, 25 + 50 kHz, SW_A field, 5 periods; bubble u's position
# ~ U[lo_u, hi_u] wavelengths.
HARNESS_SCENES = {
    "2B_TYPICAL": [(-0.45, -0.05), (0.05, 0.45)],
    "3B_TYPICAL": [(-0.45, -0.2), (-0.15, 0.15), (0.2, 0.45)],
    "4B_TYPICAL": [(-0.45, -0.3), (-0.25, -0.025), (0.025, 0.25), (0.3, 0.45)],
}


def harness_scene(scene_id: str = "2B_TYPICAL", ns: int = 4096, seed: int = 11) -> Workload:
    """A scene of the reference's kernel benchmark with its draw order
    (build_parameter_ranges, kernel_prof.py:120-150): positions bubble by bubble, then
    the amplitudes harmonic by harmonic, from numpy default_rng(seed)."""
    xr = HARNESS_SCENES[scene_id]
    n = len(xr)
    rng = np.random.default_rng(seed)
    R0 = np.full((ns, n), 60.0)
    X0 = np.empty((ns, n))
    for u, (lo, hi) in enumerate(xr):
        X0[:, u] = rng.uniform(lo, hi, size=(ns,))
    PA = np.empty((2, ns))
    for k in range(2):
        PA[k] = rng.uniform(0.0, 1.0, size=(ns,))
    FR = np.tile(np.array([[25.0], [50.0]]), (1, ns))
    return Workload(f"harness_{scene_id}_{ns}", n, R0, X0, PA, FR, np.zeros((2, ns)), np.full(ns, 5.0),
                    field="SW_A")


def build_model(w: Workload, backend=None, num_stored_points: int = 0, solver_spec=None, variant=None,
                block_size: Optional[int] = None, kernel_backend=True):
    """CKM1D model for a workload: parameters built, initial positions and horizons set
    in the host buffers (device buffers are created on the first solve)."""
    from .backends import BackendName, KernelVariant
    from .models.ckm1d import CKM1D
    from .models.specs import SolverSpec

    spec = solver_spec or SolverSpec(atol=w.atol, rtol=w.rtol)
    m = CKM1D(num_systems=w.ns, systems_per_block=64, units_per_system=w.n, R0=w.R0, PA=w.PA, FR=w.FR,
              PS=w.PS, REF_FREQ=w.ref_freq, num_components=w.nk, acoustic_field=w.field, mat_props=w.mat,
              backend=backend or BackendName.HIP, variant=variant or KernelVariant.AUTO,
              num_stored_points=num_stored_points, solver_spec=spec, block_size=block_size)
    hb = m.runtime.host_buffers
    hb.state.actual_state[1, :, :w.n] = w.X0
    hb.state.time_end[:] = w.t_end
    return m
