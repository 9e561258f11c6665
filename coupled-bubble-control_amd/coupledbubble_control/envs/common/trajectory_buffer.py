"""Per-episode trajectories of the vector environment and their zarr files
(reference: coupledbubble_control/envs/common/trajectory_buffer.py:1-230).

``TrajectorBuffer`` (the reference's spelling) keeps every env's running
episode on the device (cbc_traj_append after each step) and, when episodes
end, packs the finished ones into one buffer (cbc_traj_gather), copies it to
the host once and writes one zarr group per episode with
``TrajectoryZarrWriter``: the reference's shard directories
``shard_{start:05d}_{end:05d}.zarr`` holding groups ``trj_{id:06d}`` with the
float32 arrays observations (T+1, obs), actions (T, A), rewards (T, 1),
equilibrium_radii (N,), dense_time (P,), dense_state_{row} (P, N) and the
group attributes env_id, num_units, episode_length, episode_reward,
saved_unix; ``global_metadata.zarr`` carries the environment's metadata.
The stores are Zarr v3 (zarr_v3.py), written without the zarr package.

The merge rules are the reference's (Trajectory.merge_list, :38-56): the
observation list is the step observations followed by the final
observation, the dense points of each step after the first lose their first
point (the previous step's last) and are offset by k * time_step_length with
the Trajectory's time_step_length of 5.0 (the value the reference's
environment leaves in place).
"""
from __future__ import annotations

import ctypes
import os
import time
from collections import defaultdict
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any, DefaultDict, Dict, List, Optional, Sequence

import numpy as np
import torch

from . import zarr_v3


@dataclass
class Trajectory:
    """One episode's data (reference :15-68); lists of per-step arrays."""
    env_id: int
    num_units: int
    equilibrium_radii: Optional[np.ndarray] = None
    observations: List[np.ndarray] = field(default_factory=list)
    actions: List[np.ndarray] = field(default_factory=list)
    rewards: List[np.ndarray] = field(default_factory=list)
    dense_time: List[np.ndarray] = field(default_factory=list)
    dense_states: DefaultDict[Any, List[np.ndarray]] = field(default_factory=lambda: defaultdict(list))
    episode_length: int = 0
    episode_reward: float = 0
    time_step_length: float = 5.0

    def __post_init__(self):
        if self.equilibrium_radii is None:
            self.equilibrium_radii = np.zeros(self.num_units, dtype=np.float32)

    def merge_list(self) -> Dict[str, Any]:
        obs = np.vstack(self.observations)
        act = np.vstack(self.actions)
        rew = np.vstack(self.rewards)
        t = np.concatenate([d if i == 0 else d[1:] + i * self.time_step_length for i, d in enumerate(self.dense_time)])
        states = {k: np.concatenate([s if i == 0 else s[1:, :] for i, s in enumerate(v)])
                  for k, v in self.dense_states.items()}
        return dict(observations=obs, actions=act, rewards=rew, dense_time=t, dense_state=states)

    def clear(self) -> None:
        if self.equilibrium_radii is not None:
            self.equilibrium_radii.fill(0.0)
        self.observations.clear()
        self.actions.clear()
        self.rewards.clear()
        self.dense_time.clear()
        self.dense_states.clear()
        self.episode_length = 0
        self.episode_reward = 0.0


class TrajectoryZarrWriter:
    """One zarr group per episode in shards of ``shard_size`` episodes (reference :158-226)."""

    def __init__(self, root_dir: str, shard_size: int = 1000, global_metadata: Optional[Dict] = None,
                 verbose: bool = True, async_write: bool = False) -> None:
        """``async_write`` writes the episodes in order on one background thread, so
        file output overlaps the next steps; ``flush()`` / ``close()`` wait for it."""
        self.root_dir = Path(root_dir)
        self.root_dir.mkdir(parents=True, exist_ok=True)
        if verbose:
            print("Trajectory Dir", self.root_dir)
        self.shard_size = shard_size
        self.global_metadata: Dict = global_metadata if global_metadata is not None else {}
        self._write_global_metadata()
        self._shards = set()
        self._queue = self._thread = None
        self._error: Optional[BaseException] = None
        if async_write:
            import queue
            import threading

            self._queue = queue.Queue()
            self._thread = threading.Thread(target=self._run, name="trajectory-writer", daemon=True)
            self._thread.start()

    def _write_global_metadata(self):
        if not self.global_metadata:
            return
        path = self.root_dir / "global_metadata.zarr"
        zarr_v3.open_group(path)
        if not zarr_v3.group_attrs(path).get("_initialized", False):
            zarr_v3.update_attrs(path, self.global_metadata)
            zarr_v3.update_attrs(path, {"_initialized": True})

    def _shard_path(self, episode_id: int) -> Path:
        start = (episode_id // self.shard_size) * self.shard_size
        return self.root_dir / f"shard_{start:05d}_{start + self.shard_size - 1:05d}.zarr"

    def _group_path(self, episode_id: int) -> Path:
        """The episode's group path in its shard (the shard created when missing, an
        existing group of the same name removed, as the reference does)."""
        shard = self._shard_path(episode_id)
        if shard not in self._shards:
            zarr_v3.open_group(shard)
            self._shards.add(shard)
        g = shard / f"trj_{episode_id:06d}"
        if os.path.lexists(g):
            zarr_v3.delete(g)
        return g

    def _write(self, episode_id: int, trj: Trajectory) -> None:
        data = trj.merge_list()
        arrays = {"observations": data["observations"], "actions": data["actions"], "rewards": data["rewards"]}
        if trj.equilibrium_radii is not None:
            arrays["equilibrium_radii"] = trj.equilibrium_radii
        arrays["dense_time"] = data["dense_time"]
        for idx, arr in data["dense_state"].items():
            arrays[f"dense_state_{idx:0d}"] = arr
        zarr_v3.write_new_group(self._group_path(episode_id), arrays, dict(
            env_id=int(episode_id), num_units=int(trj.num_units), episode_length=trj.episode_length,
            episode_reward=trj.episode_reward, saved_unix=int(time.time())))

    def write(self, episode_id: int, trj: Trajectory) -> None:
        """One episode's group (reference :196-226); queued when writing in the background."""
        if self._queue is not None:
            if self._error is not None:
                raise RuntimeError("trajectory writer thread failed") from self._error
            self._queue.put((episode_id, trj))
            return
        self._write(episode_id, trj)

    def flush(self) -> None:
        """Wait until every queued episode is on disk (background writing only)."""
        if self._queue is not None:
            self._queue.join()
            if self._error is not None:
                raise RuntimeError("trajectory writer thread failed") from self._error

    def close(self) -> None:
        if self._queue is not None:
            self.flush()
            self._queue.put(None)
            self._thread.join()
            self._queue = None

    def _run(self) -> None:
        while True:
            item = self._queue.get()
            try:
                if item is None:
                    return
                if self._error is None:
                    self._write(*item)
            except BaseException as e:  # reported to the caller on its next write / flush
                self._error = e
            finally:
                self._queue.task_done()


class TrajectorBuffer:
    """Device-resident episode collection (reference :70-156); same methods.

    ``max_episode_steps`` bounds the steps one episode may record (the arena is
    allocated at the first step, from the shapes of its inputs); a longer
    episode raises at its end."""

    def __init__(self, num_envs: int, num_units: int, saved_states: Optional[Sequence[int]] = None, *,
                 zarr_root: str, zarr_shard_size: int = 1000, metadata: Optional[Dict] = None,
                 max_episode_steps: int = 128, time_step_length: float = 5.0, verbose: bool = True,
                 async_write: bool = False) -> None:
        self.episode_count = 0
        self._num_envs = num_envs
        self._num_units = num_units
        self._state_index: List[int] = list(saved_states) if saved_states is not None else [0, 1]
        self._max_steps = int(max_episode_steps)
        self._step_length = float(time_step_length)
        self._verbose = verbose
        self._zarr_writer = TrajectoryZarrWriter(zarr_root, zarr_shard_size, metadata, verbose=verbose,
                                                 async_write=async_write)
        self._arena = None

    def flush(self) -> None:
        """Wait for the background writer (no-op for synchronous writing)."""
        self._zarr_writer.flush()

    def close(self) -> None:
        self._zarr_writer.close()

    # -- the device arena ------------------------------------------------------
    def _allocate(self, observations, actions, dense_state):
        from ...backends.hip import _abi

        ndo, sd, ne, tile = dense_state.shape
        if ne != self._num_envs:
            raise ValueError(f"dense_state holds {ne} envs, the buffer {self._num_envs}")
        dev = dense_state.device
        s = _abi.TrajSpecC()
        s.num_envs, s.num_bubbles, s.tile, s.ndo = ne, self._num_units, tile, ndo
        s.num_saved = len(self._state_index)
        for j, r in enumerate(self._state_index):
            s.saved_rows[j] = int(r)
        s.obs_dim = int(observations.shape[1])
        s.num_actions = int(actions.shape[1])
        s.max_steps = self._max_steps
        s.max_points = self._max_steps * ndo
        s.step_length = self._step_length
        f32 = dict(dtype=torch.float32, device=dev)
        a = {
            "obs": torch.zeros((ne, s.max_steps, s.obs_dim), **f32),
            "act": torch.zeros((ne, s.max_steps, max(1, s.num_actions)), **f32),
            "rew": torch.zeros((ne, s.max_steps), **f32),
            "time": torch.zeros((ne, s.max_points), **f32),
            "states": torch.zeros((ne, max(1, s.num_saved), s.max_points, self._num_units), **f32),
            "steps": torch.zeros(ne, dtype=torch.int32, device=dev),
            "points": torch.zeros(ne, dtype=torch.int64, device=dev),
            "overflow": torch.zeros(1, dtype=torch.int32, device=dev),
        }
        self._spec, self._arena = s, a

    def _buffers(self, step_inputs=None):
        from ...backends.hip import _abi

        b = _abi.TrajBuffersC()
        for k in ("obs", "act", "rew", "time", "states", "steps", "points", "overflow"):
            setattr(b, k, self._arena[k].data_ptr())
        if step_inputs is not None:
            for k, t in step_inputs.items():
                setattr(b, k, t.data_ptr())
        return b

    def step(self, observations: torch.Tensor, actions: torch.Tensor, rewards: torch.Tensor,
             dense_index: torch.Tensor, dense_time: torch.Tensor, dense_states: torch.Tensor) -> None:
        """Record one step of every env (reference :95-127); clears dense_index."""
        from ...backends.hip import _abi

        if self._arena is None:
            self._allocate(observations, actions, dense_states)
        dev = dense_states.device
        inputs = {
            "observations": observations.to(device=dev, dtype=torch.float32).contiguous(),
            "actions": actions.to(device=dev, dtype=torch.float32).reshape(self._num_envs, -1).contiguous(),
            "rewards": rewards.to(device=dev, dtype=torch.float32).reshape(-1).contiguous(),
            "dense_index": dense_index, "dense_time": dense_time, "dense_state": dense_states,
        }
        if dense_index.dtype != torch.int32 or dense_time.dtype != torch.float64 or dense_states.dtype != torch.float64:
            raise TypeError("dense buffers must be int32 / float64 device tensors")
        self._keep = inputs
        b = self._buffers(inputs)
        _abi.check(_abi.lib().cbc_traj_append(ctypes.byref(self._spec), ctypes.byref(b), _abi.stream_handle()))

    def end_episode(self, env_ids, final_observations: torch.Tensor, episode_lengths, episode_rewards,
                    equilibrium_radii) -> None:
        """Write the finished episodes of ``env_ids`` (reference :129-155)."""
        from ...backends.hip import _abi

        if self._verbose:
            print("Trj buffer end episode")
        ids = torch.as_tensor(env_ids, dtype=torch.int64).reshape(-1)
        n = int(ids.numel())
        if n == 0 or self._arena is None:
            return
        if int(self._arena["overflow"].item()):
            raise RuntimeError(f"an episode recorded more than max_episode_steps = {self._max_steps} steps")
        dev = self._arena["steps"].device
        ids_d = ids.to(dev)
        steps = self._arena["steps"][ids_d].to(torch.int64).cpu().numpy()
        points = self._arena["points"][ids_d].cpu().numpy()
        step_off = np.concatenate([[0], np.cumsum(steps)[:-1]]).astype(np.int64)
        point_off = np.concatenate([[0], np.cumsum(points)[:-1]]).astype(np.int64)
        S, P = int(steps.sum()), int(points.sum())
        s = self._spec
        ns = max(1, s.num_saved)
        out = {"obs": torch.empty((max(S, 1), s.obs_dim), dtype=torch.float32, device=dev),
               "act": torch.empty((max(S, 1), max(1, s.num_actions)), dtype=torch.float32, device=dev),
               "rew": torch.empty(max(S, 1), dtype=torch.float32, device=dev),
               "time": torch.empty(max(P, 1), dtype=torch.float32, device=dev),
               "states": torch.empty((ns, max(P, 1), self._num_units), dtype=torch.float32, device=dev)}
        pk = _abi.TrajPackedC()
        for k, t in out.items():
            setattr(pk, k, t.data_ptr())
        pk.total_points = P
        so = torch.from_numpy(step_off).to(dev)
        po = torch.from_numpy(point_off).to(dev)
        _abi.check(_abi.lib().cbc_traj_gather(ctypes.byref(s), ctypes.byref(self._buffers()), ids_d.data_ptr(), n,
                                              so.data_ptr(), po.data_ptr(), ctypes.byref(pk), _abi.stream_handle()))
        host = {k: v.cpu().numpy() for k, v in out.items()}
        states = host["states"] if P else np.zeros((ns, 0, self._num_units), dtype=np.float32)
        final_obs = final_observations.cpu().numpy()
        ep_len = torch.as_tensor(episode_lengths).cpu().numpy()
        ep_rew = torch.as_tensor(episode_rewards).cpu().numpy()
        eq = np.asanyarray(equilibrium_radii, dtype=np.float32)
        ids_h = ids.cpu().numpy()
        for i in range(n):
            a0, a1 = int(step_off[i]), int(step_off[i] + steps[i])
            p0, p1 = int(point_off[i]), int(point_off[i] + points[i])
            trj = Trajectory(env_id=int(ids_h[i]), num_units=self._num_units, time_step_length=self._step_length)
            trj.observations = [host["obs"][a0:a1], final_obs[i][None, :]]
            trj.actions = [host["act"][a0:a1, :s.num_actions]]
            trj.rewards = [host["rew"][a0:a1].reshape(-1, 1)]
            trj.dense_time = [host["time"][p0:p1]]
            for j, r in enumerate(self._state_index):
                trj.dense_states[r] = [states[j, p0:p1]]
            trj.episode_length = ep_len[i]
            trj.episode_reward = ep_rew[i]
            trj.equilibrium_radii = eq[i]
            self._zarr_writer.write(self.episode_count, trj)
            self.episode_count += 1
