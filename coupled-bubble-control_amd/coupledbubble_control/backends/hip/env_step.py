"""Host side of the environment kernels (include/cbc.h: cbc_env_*).

``EnvKernels`` owns the environment's device buffers next to one model's
solver buffers -- the observation history (one float32 array holding every
observation part), the episode counters, the done list and the kernels'
workspace -- and calls the three entry points that replace the torch code
of PosNBC1D.step (coupledbubble_control/envs/bubble_1D_pos_control.py:731-853).
Per-step outputs (observations, rewards, flags) are fresh tensors each step,
as the reference's are, so tensors a caller keeps stay valid.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Optional, Sequence

import torch

from . import _abi


@dataclass
class ObsPart:
    """One observation component: kind (_abi.CBC_ENV_PART_*), state row, ring depth."""
    name: str
    kind: int
    row: int
    stack: int


@dataclass
class StepOutputs:
    observations: torch.Tensor   # (NE, obs_dim) float32
    rewards: torch.Tensor        # (NE,) float32
    failure: torch.Tensor        # (NE,) bool
    success: torch.Tensor
    terminated: torch.Tensor
    time_outs: torch.Tensor


class EnvKernels:
    def __init__(self, model, parts: Sequence[ObsPart], obs_bias: torch.Tensor, obs_rscale: torch.Tensor, *,
                 action_rows: Sequence[int], action_scales: Sequence[float], pa_cols: Sequence[int],
                 position_tolerance: float, dnorm: float, reward_exp: float, weights: Sequence[float],
                 positive_terminal_reward: float, negative_terminal_reward: float, apply_termination: bool,
                 episode_length: int, next_time_step: float = 1e-6):
        es = model._execution_spec
        self.model = model
        self.ne, self.n, self.tile = es.ns, es.ups, es.tile
        if self.n > _abi.CBC_ENV_MAX_BUBBLES:
            raise ValueError(f"the environment kernels take at most {_abi.CBC_ENV_MAX_BUBBLES} bubbles, got {self.n}")
        if len(parts) > _abi.CBC_ENV_MAX_PARTS:
            raise ValueError(f"at most {_abi.CBC_ENV_MAX_PARTS} observation components, got {len(parts)}")
        if len(action_rows) > _abi.CBC_ENV_MAX_ACTIONS:
            raise ValueError(f"at most {_abi.CBC_ENV_MAX_ACTIONS} action columns, got {len(action_rows)}")
        model._ensure_buffers()
        db = model.runtime.device_buffers
        dev = db.state.actual_state.device
        self.device = dev
        self.parts: List[ObsPart] = list(parts)
        s = _abi.EnvSpecC()
        s.num_envs, s.num_bubbles, s.tile = self.ne, self.n, self.tile
        s.num_control_rows = int(db.params.control_params.shape[0])
        s.num_actions = len(action_rows)
        for c, (r, sc) in enumerate(zip(action_rows, action_scales)):
            s.action_row[c] = int(r)
            s.action_scale[c] = float(sc)
        s.num_pa = len(pa_cols)
        for c, col in enumerate(pa_cols):
            s.pa_col[c] = int(col)
        s.num_parts = len(self.parts)
        off, dim = 0, 0
        self.part_offset = []
        for p_i, p in enumerate(self.parts):
            s.part_kind[p_i], s.part_row[p_i], s.part_stack[p_i], s.part_hist[p_i] = p.kind, p.row, p.stack, off
            self.part_offset.append(off)
            off += p.stack * self.ne * self.n
            dim += p.stack * self.n
        s.obs_dim = dim
        s.history_floats = off
        names = [p.name for p in self.parts]
        s.x_part = names.index("X")
        s.target_part = names.index("XT")
        s.apply_termination = int(bool(apply_termination))
        s.episode_length = int(episode_length)
        s.position_tolerance = float(position_tolerance)
        s.r_dnorm = 1.0 / float(dnorm)
        s.reward_exp = float(reward_exp)
        s.w_distance, s.w_intensity = float(weights[0]), float(weights[1])
        s.positive_terminal_reward = float(positive_terminal_reward)
        s.negative_terminal_reward = float(negative_terminal_reward)
        s.next_time_step = float(next_time_step)
        self.spec = s
        self.obs_dim = dim
        if tuple(obs_bias.shape) != (dim,) or tuple(obs_rscale.shape) != (dim,):
            raise ValueError(f"observation bounds must have {dim} columns")
        self.obs_bias = obs_bias.to(device=dev, dtype=torch.float32).contiguous()
        self.obs_rscale = obs_rscale.to(device=dev, dtype=torch.float32).contiguous()
        self.history = torch.zeros(max(off, 1), dtype=torch.float32, device=dev)
        self.total_rewards = torch.zeros(self.ne, dtype=torch.float32, device=dev)
        self.algo_steps = torch.zeros(self.ne, dtype=torch.int64, device=dev)
        self.done_ids = torch.zeros(self.ne, dtype=torch.int64, device=dev)
        words = int(_abi.lib().cbc_env_workspace_words(self.ne))
        self.workspace = torch.zeros(words, dtype=torch.int32, device=dev)
        self.actions = torch.zeros((self.ne, max(1, s.num_actions)), dtype=torch.float32, device=dev)
        self.observations = torch.zeros((self.ne, dim), dtype=torch.float32, device=dev)
        self.outputs: Optional[StepOutputs] = None

    # ------------------------------------------------------------------
    def part_view(self, name: str) -> torch.Tensor:
        """History of one part: (stack, NE, N) for a state row or a stacked part, (NE, N) otherwise."""
        i = [p.name for p in self.parts].index(name)
        p = self.parts[i]
        v = self.history[self.part_offset[i]:self.part_offset[i] + p.stack * self.ne * self.n]
        if p.kind == _abi.CBC_ENV_PART_STATE or p.stack > 1:
            return v.view(p.stack, self.ne, self.n)
        return v.view(self.ne, self.n)

    def _buffers(self, out: Optional[StepOutputs]) -> _abi.EnvBuffersC:
        db = self.model.runtime.device_buffers
        p = _abi.ptr
        b = _abi.EnvBuffersC()
        b.actual_state, b.actual_time = p(db.state.actual_state), p(db.state.actual_time)
        b.time_step, b.actual_event = p(db.state.time_step), p(db.status.actual_event)
        b.status_flags, b.control_params = p(db.status.status_flags), p(db.params.control_params)
        b.actions, b.history = p(self.actions), p(self.history)
        b.observations = p(out.observations if out is not None else self.observations)
        b.obs_bias, b.obs_rscale = p(self.obs_bias), p(self.obs_rscale)
        b.total_rewards, b.algo_steps = p(self.total_rewards), p(self.algo_steps)
        b.done_ids, b.workspace = p(self.done_ids), p(self.workspace)
        if out is not None:
            b.rewards, b.failure, b.success = p(out.rewards), p(out.failure), p(out.success)
            b.terminated, b.time_outs = p(out.terminated), p(out.time_outs)
        return b

    def _check_solver_buffers(self):
        db = self.model.runtime.device_buffers
        if db.params.control_params.dtype != torch.float64 or db.state.actual_state.dtype != torch.float64:
            raise TypeError("solver buffers must be float64")

    # ------------------------------------------------------------------
    def set_action(self, actions: torch.Tensor, stream=None) -> None:
        """actions (NE, A) -> control_params rows (set_action, :732-749)."""
        a = torch.as_tensor(actions)
        if tuple(a.shape) != (self.ne, self.spec.num_actions):
            raise ValueError(f"actions must have shape {(self.ne, self.spec.num_actions)}, got {tuple(a.shape)}")
        self.actions.copy_(a.to(device=self.device, dtype=torch.float32))
        self._check_solver_buffers()
        b = self._buffers(None)
        _abi.check(_abi.lib().cbc_env_set_action(ctypes.byref(self.spec), ctypes.byref(b),
                                                 _abi.stream_handle(stream)))

    def post(self, observed: int, stream=None, static_workspace: bool = False) -> _abi.EnvCountsC:
        """observe .. advance_time (:756-827) into fresh output tensors; returns the
        counts (synchronises the stream, as the reference's step does).  With
        ``static_workspace`` the call goes through cbc_env_post_static, which uses the
        library's own device workspace instead of this object's (same results)."""
        dev = self.device
        out = StepOutputs(
            observations=torch.empty((self.ne, self.obs_dim), dtype=torch.float32, device=dev),
            rewards=torch.empty(self.ne, dtype=torch.float32, device=dev),
            failure=torch.empty(self.ne, dtype=torch.bool, device=dev),
            success=torch.empty(self.ne, dtype=torch.bool, device=dev),
            terminated=torch.empty(self.ne, dtype=torch.bool, device=dev),
            time_outs=torch.empty(self.ne, dtype=torch.bool, device=dev))
        self._check_solver_buffers()
        b = self._buffers(out)
        counts = _abi.EnvCountsC()
        fn = _abi.lib().cbc_env_post_static if static_workspace else _abi.lib().cbc_env_post
        _abi.check(fn(ctypes.byref(self.spec), ctypes.byref(b), int(observed), _abi.stream_handle(stream),
                      ctypes.byref(counts)))
        self.outputs = out
        self.observations = out.observations
        return counts

    def _range_spec(self, begin: int, count: int) -> _abi.EnvSpecC:
        if not (0 <= begin and count >= 1 and begin + count <= self.ne):
            raise ValueError(f"env range [{begin}, {begin + count}) outside [0, {self.ne})")
        s = _abi.EnvSpecC.from_buffer_copy(self.spec)
        s.env_begin, s.env_count = int(begin), int(count)
        return s

    def set_action_range(self, begin: int, actions: torch.Tensor, stream=None) -> None:
        """set_action for envs [begin, begin + len(actions)) only (asynchronous)."""
        a = torch.as_tensor(actions)
        count = int(a.shape[0])
        if tuple(a.shape) != (count, self.spec.num_actions):
            raise ValueError(f"actions must have shape (n, {self.spec.num_actions}), got {tuple(a.shape)}")
        spec = self._range_spec(begin, count)
        self.actions[begin:begin + count].copy_(a.to(device=self.device, dtype=torch.float32))
        self._check_solver_buffers()
        b = self._buffers(None)
        _abi.check(_abi.lib().cbc_env_set_action(ctypes.byref(spec), ctypes.byref(b), _abi.stream_handle(stream)))

    def new_outputs(self) -> StepOutputs:
        """Output tensors for every env (written row range by row range by post_range)."""
        dev = self.device
        return StepOutputs(
            observations=torch.zeros((self.ne, self.obs_dim), dtype=torch.float32, device=dev),
            rewards=torch.zeros(self.ne, dtype=torch.float32, device=dev),
            failure=torch.zeros(self.ne, dtype=torch.bool, device=dev),
            success=torch.zeros(self.ne, dtype=torch.bool, device=dev),
            terminated=torch.zeros(self.ne, dtype=torch.bool, device=dev),
            time_outs=torch.zeros(self.ne, dtype=torch.bool, device=dev))

    def post_range(self, begin: int, count: int, observed: int, out: StepOutputs, stream=None) -> _abi.EnvCountsC:
        """observe .. advance_time for envs [begin, begin + count) into the rows of ``out``
        (``observed``: the posts this range has had); the done list holds the range's
        done envs (global ids, ascending).  Returns the counts (synchronises the stream)."""
        spec = self._range_spec(begin, count)
        self._check_solver_buffers()
        b = self._buffers(out)
        counts = _abi.EnvCountsC()
        _abi.check(_abi.lib().cbc_env_post(ctypes.byref(spec), ctypes.byref(b), int(observed),
                                           _abi.stream_handle(stream), ctypes.byref(counts)))
        return counts

    def reset(self, scene: _abi.EnvSceneC, ids: torch.Tensor, observed: int, final: bool, stream=None,
              observations: Optional[torch.Tensor] = None):
        """Final outputs of ``ids`` (when ``final``) then their reset (:583-614,
        :833-840).  Observation rows are rebuilt in ``observations`` (default:
        the current observation tensor)."""
        ids = ids.to(device=self.device, dtype=torch.int64).contiguous()
        n = int(ids.numel())
        if observations is not None:
            self.observations = observations
        fo = fr = fl = None
        if final:
            fo = torch.empty((n, self.obs_dim), dtype=torch.float32, device=self.device)
            fr = torch.empty(n, dtype=torch.float32, device=self.device)
            fl = torch.empty(n, dtype=torch.int64, device=self.device)
        b = self._buffers(None)
        _abi.check(_abi.lib().cbc_env_reset(ctypes.byref(self.spec), ctypes.byref(b), ctypes.byref(scene),
                                            _abi.ptr(ids), n, int(observed), _abi.ptr(fo), _abi.ptr(fr),
                                            _abi.ptr(fl), _abi.stream_handle(stream)))
        return fo, fr, fl


def scene_positions(scene: _abi.EnvSceneC, n: int, num_bubbles: int, device, stream=None):
    """(positions, targets) float64 (n, N) of n scenes on ``device``."""
    pos = torch.empty((n, num_bubbles), dtype=torch.float64, device=device)
    tgt = torch.empty((n, num_bubbles), dtype=torch.float64, device=device)
    _abi.check(_abi.lib().cbc_env_scene_positions(ctypes.byref(scene), int(n), int(num_bubbles), _abi.ptr(pos),
                                                  _abi.ptr(tgt), _abi.stream_handle(stream)))
    return pos, tgt
