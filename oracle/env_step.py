"""CPU restatement of the vector environment's step around the solver.  TEST INFRASTRUCTURE ONLY.

Restates in NumPy what PosNBC1D does on the solver buffers between two
solves (coupledbubble_control/envs/bubble_1D_pos_control.py):

  set_action          :732-749   actions * scale into control_params rows
  observe             :756-769   state rows -> float32 observation buffers
  check_termination   :772-789   failure = status != 1, status := 0, success
  get_rewards         :792-813   distance and actuation penalties, terminal rewards
  advance_time        :816-827   t := 0, h := 1e-6, event := -1, step counters
  final_termination   :830-853   done list and final outputs, then reset_envs
  reset_envs          :583-614   new scenes written into state and buffers
  SceneGenerator      :46-249    random / equidistant / fixed positions

with the observation buffers of envs/common/buffers.py (RingBuffer with its
shared pointer, PersistentBuffer, get_stacked) and the Box scaling of
envs/common/spaces.py:84-117 / specs.py:29-44.

Arithmetic: the reference evaluates observations and rewards in float32, one
elementwise torch kernel per operation; here every such operation is a
float32 NumPy operation.  Sums over bubbles and gaps are taken in index order
(cumulative sums; NumPy's pairwise summation would reorder them) and
torch.mean is sum * float32(1/N), the form of ATen's mean reduction.  The
order in which torch's reduction kernels add a row of 2..16 values is not
specified, so the reference may differ from this restatement by an ulp in a
sum of more than two terms; the tests state their tolerance accordingly.

Pinning (tests/test_oracle_env.py): hand-derived known answers for rewards,
terminations and counters, the ring order and the repeated-bounds scaling of
the observation space, closed forms of the scene generator (zero draws, the
clamped-gap limit, alignment extremes, symmetric centred scenes); evenly
spaced scenes without a distance take torch.linspace itself.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

F32 = np.float32


# ---------------------------------------------------------------------------
# observation buffers (envs/common/buffers.py)

class Persistent:
    """PersistentBuffer (buffers.py:21-33): one (NE, count) float32 array."""

    def __init__(self, ne: int, count: int):
        self.buffer = np.zeros((ne, count), dtype=F32)

    def update(self, x, fill: bool = False):
        self.buffer[:] = np.asarray(x).astype(F32)

    def update_at(self, ids, x, fill: bool = False):
        self.buffer[ids] = np.asarray(x).astype(F32)

    def get(self):
        return self.buffer

    def get_stacked(self):
        return self.buffer


class Ring:
    """RingBuffer (buffers.py:36-83): (capacity, NE, count) with one shared pointer."""

    def __init__(self, capacity: int, ne: int, count: int):
        self.buffer = np.zeros((capacity, ne, count), dtype=F32)
        self.capacity = capacity
        self.ptr = 0

    def update(self, x, fill: bool = False):
        x = np.asarray(x).astype(F32)
        if fill:
            self.buffer[:, :] = x
        else:
            self.buffer[self.ptr] = x
            self.ptr = (self.ptr + 1) % self.capacity

    def update_at(self, ids, x, fill: bool = False):
        x = np.asarray(x).astype(F32)
        if fill:
            self.buffer[:, ids] = x
        else:
            self.buffer[self.ptr, ids] = x

    def get(self):
        return self.buffer[(self.ptr - 1) % self.capacity]

    def get_stacked(self):
        idx = np.arange(self.ptr, self.ptr + self.capacity) % self.capacity
        st = self.buffer[idx].transpose(1, 2, 0)          # (NE, count, capacity)
        return st.reshape(st.shape[0], -1)


@dataclass
class Part:
    """One observation component (ObservationComponent, specs.py:20-28)."""
    name: str
    low: List[float]
    high: List[float]
    stack: int = 1


def box_scale(parts: Sequence[Part]):
    """(bias, rscale) float32 of SpaceSchema.build_space -> Box (specs.py:35-44,
    spaces.py:84-113): each component's bounds are its list repeated `stack`
    times (list repetition), concatenated over components."""
    low = np.concatenate([np.asarray(list(p.low) * p.stack, dtype=F32) for p in parts])
    high = np.concatenate([np.asarray(list(p.high) * p.stack, dtype=F32) for p in parts])
    scale = (high - low) * F32(0.5)
    rscale = F32(1.0) / scale
    bias = (high + low) * F32(0.5)
    return bias, rscale


def pow_scalar(x: np.ndarray, e: float) -> np.ndarray:
    """torch.pow(float32 tensor, python scalar): ATen's dedicated formulas for
    0.5, 2, 3, -0.5, -1, -2, powf otherwise."""
    x = x.astype(F32)
    if e == 0.5:
        return np.sqrt(x)
    if e == 1.0:
        return x
    if e == 2.0:
        return x * x
    if e == 3.0:
        return x * x * x
    if e == -0.5:
        return F32(1.0) / np.sqrt(x)
    if e == -1.0:
        return F32(1.0) / x
    if e == -2.0:
        return F32(1.0) / (x * x)
    return np.power(x, F32(e))


def _row_sum(x: np.ndarray) -> np.ndarray:
    """Sum along axis 1 in index order, in the array's dtype."""
    if x.shape[1] == 0:
        return np.zeros(x.shape[0], dtype=x.dtype)
    return np.cumsum(x, axis=1, dtype=x.dtype)[:, -1]


# ---------------------------------------------------------------------------
# the environment's step logic

STATE_ROWS = {"R": 0, "X": 1, "U": 2, "V": 3}


class EnvOracle:
    """The per-step logic of PosNBC1D on host arrays (the solver buffers are
    passed in: state (4, NE, TILE), time/step/event (NE,), status (NE,) int32,
    control_params (4 NK, NE))."""

    def __init__(self, num_envs: int, num_bubbles: int, parts: Sequence[Part], *, action_rows: Sequence[int],
                 action_scales: Sequence[float], pa_cols: Sequence[int], position_tolerance: float = 0.01,
                 dnorm: float = 0.2, reward_exp: float = 0.5, weights: Sequence[float] = (1.0, 0.0),
                 positive_terminal_reward: float = 10.0, negative_terminal_reward: float = -10.0,
                 apply_termination: bool = True, episode_length: int = 128, next_time_step: float = 1e-6):
        self.ne, self.n = num_envs, num_bubbles
        self.parts = list(parts)
        self.buffers: Dict[str, object] = {}
        for p in self.parts:
            self.buffers[p.name] = (Ring(p.stack, num_envs, num_bubbles) if p.stack > 1
                                    else Persistent(num_envs, num_bubbles))
        self.bias, self.rscale = box_scale(self.parts)
        self.action_rows = list(action_rows)
        self.action_scales = np.asarray(action_scales, dtype=np.float64)
        self.pa_cols = list(pa_cols)
        self.tol = F32(position_tolerance)
        self.r_dnorm = F32(1.0 / dnorm)
        self.reward_exp = float(reward_exp)
        self.weights = np.asarray(weights, dtype=F32)
        self.pos_r = F32(positive_terminal_reward)
        self.neg_r = F32(negative_terminal_reward)
        self.apply_termination = apply_termination
        self.episode_length = int(episode_length)
        self.next_time_step = next_time_step
        self.algo_steps = np.zeros(num_envs, dtype=np.int64)
        self.total_rewards = np.zeros(num_envs, dtype=F32)

    # _get_observation (:725-729)
    def observations(self) -> np.ndarray:
        stacked = np.concatenate([self.buffers[p.name].get_stacked() for p in self.parts], axis=1)
        return (stacked - self.bias) * self.rscale

    # set_action (:732-749)
    def set_action(self, actions: np.ndarray, control_params: np.ndarray) -> None:
        a64 = np.asarray(actions).astype(np.float64) * self.action_scales
        control_params[self.action_rows, :] = a64.T

    def post(self, state, actual_time, time_step, actual_event, status, actions) -> dict:
        """observe, check_termination, get_rewards, advance_time and the done list
        of final_termination (:756-832); mutates the solver arrays like the reference."""
        n = self.n
        # observe (:756-769)
        for p in self.parts:
            if p.name in STATE_ROWS:
                self.buffers[p.name].update(state[STATE_ROWS[p.name], :, :n].astype(F32))
        obs = self.observations()
        # check_termination (:772-789)
        failure = status != 1
        status[:] = 0
        td = np.abs(self.buffers["XT"].get() - self.buffers["X"].get())
        success = np.all(td <= self.tol, axis=1) if self.apply_termination else np.zeros(self.ne, dtype=bool)
        terminated = failure | success
        # get_rewards (:792-813)
        tdm = np.where(td < self.tol, F32(0.0), td).astype(F32)
        nd = tdm * self.r_dnorm
        nd = np.where(nd < F32(0.0), F32(0.0), nd).astype(F32)
        distance = _row_sum(pow_scalar(nd, self.reward_exp)) * F32(1.0 / n)
        a = np.asarray(actions, dtype=F32)[:, self.pa_cols]
        intensity = np.sqrt(_row_sum(a * a))
        rewards = (-self.weights[0]) * distance
        rewards = rewards - self.weights[1] * intensity
        rewards = rewards + self.neg_r * failure.astype(F32)
        rewards = rewards + self.pos_r * success.astype(F32)
        self.total_rewards = self.total_rewards + rewards
        # advance_time (:816-827)
        actual_time[:] = 0.0
        if self.next_time_step > 0:
            time_step[:] = self.next_time_step
        actual_event[:] = -1.0
        self.algo_steps = self.algo_steps + 1
        time_outs = self.algo_steps >= self.episode_length
        done = np.flatnonzero(time_outs | terminated)
        return {"observations": obs, "rewards": rewards.astype(F32), "failure": failure, "success": success,
                "terminated": terminated, "time_outs": time_outs, "done_ids": done}

    def reset_envs(self, ids: np.ndarray, positions: np.ndarray, targets: np.ndarray, state: np.ndarray) -> dict:
        """final_termination's outputs (:833-838) then reset_envs (:583-614)."""
        n = self.n
        final = {"final_observation": self.observations()[ids], "episode_return": self.total_rewards[ids].copy(),
                 "episode_length": self.algo_steps[ids].copy()}
        state[0, ids, :n] = 1.0
        state[1, ids, :n] = positions
        state[2, ids, :n] = 0.0
        state[3, ids, :n] = 0.0
        self.buffers["XT"].update_at(ids, targets)
        self.buffers["X"].update_at(ids, positions, fill=True)
        self.algo_steps[ids] = 0
        self.total_rewards[ids] = F32(0.0)
        final["observations"] = self.observations()
        return final


# ---------------------------------------------------------------------------
# SceneGenerator (:46-249)

@dataclass
class SceneParams:
    num_bubbles: int
    domain_min: float
    domain_max: float
    margin: float = 0.05
    min_distance: float = 0.05
    max_distance: float = 0.9
    alignment: str = "center"

    @property
    def lo(self) -> float:
        return self.domain_min + self.margin

    @property
    def hi(self) -> float:
        return self.domain_max - self.margin


def random_uniform_positions(sp: SceneParams, u_gaps: np.ndarray, u_align: Optional[np.ndarray] = None) -> np.ndarray:
    """_random_uniform_position_1D (:183-242) for draws u_gaps (n, N-1) and, for
    random alignment, u_align (n,)."""
    lo, hi = sp.lo, sp.hi
    available = hi - lo
    num_gaps = sp.num_bubbles - 1
    min_required = num_gaps * sp.min_distance
    if min_required > available:
        raise ValueError("scene does not fit")
    room = sp.max_distance - sp.min_distance
    extra_limit = available - min_required
    u = np.asarray(u_gaps, dtype=np.float64)
    if u.ndim != 2:
        u = u.reshape(-1, num_gaps)
    extras_raw = u * room
    s = _row_sum(extras_raw)
    denom = np.where(s < 1e-12, 1e-12, s)
    scale = extra_limit / denom
    scale = np.where(scale > 1.0, 1.0, scale)
    gaps = extras_raw * scale[:, None]
    gaps = gaps + sp.min_distance
    remaining = available - _row_sum(gaps)
    if sp.alignment == "center":
        offset = lo + 0.5 * remaining
    else:
        offset = lo + np.asarray(u_align, dtype=np.float64).reshape(-1) * remaining
    csum = np.cumsum(gaps, axis=1) if num_gaps else np.zeros((gaps.shape[0], 0))
    return np.concatenate([offset[:, None], offset[:, None] + csum], axis=1)


def torch_linspace(lo: float, hi: float, n: int) -> np.ndarray:
    """torch.linspace(lo, hi, n) in float64 on the CPU, the library routine the
    reference calls (its rounding depends on the CPU's vector width, so it is
    evaluated, not restated)."""
    import torch

    return torch.linspace(lo, hi, n, dtype=torch.float64).numpy()


def equidistant_positions(sp: SceneParams, distance: Optional[float], n_scenes: int,
                          u_align: Optional[np.ndarray] = None) -> np.ndarray:
    """_equidistant_position_1D (:139-181)."""
    lo, hi = sp.lo, sp.hi
    N = sp.num_bubbles
    if distance is None:
        base = torch_linspace(lo, hi, N)
        gap = base[1] - base[0]
        if gap < sp.min_distance or gap > sp.max_distance:
            raise ValueError("gap out of range")
        offset = np.zeros(n_scenes)
    else:
        span = (N - 1) * distance
        if span > hi - lo:
            raise ValueError("scene does not fit")
        base = np.arange(N, dtype=np.float64) * distance
        if sp.alignment == "center":
            offset = np.full(n_scenes, (lo + hi - span) * 0.5)
        else:
            max_start = hi - span
            offset = lo + np.asarray(u_align, dtype=np.float64).reshape(-1) * (max_start - lo)
    return base[None, :] + offset[:, None]
