"""CPU restatement of the environment's trajectory recorder.  TEST INFRASTRUCTURE ONLY.

Restates TrajectorBuffer.step / end_episode and Trajectory.merge_list
(coupledbubble_control/envs/common/trajectory_buffer.py:15-155) on host
arrays, list-based as the reference is: every step appends each env's
observation, action, reward and its dense points [0, dense_index) of the saved
state rows, then dense_index is cleared; an episode's end appends the final
observation and merges the lists -- the first point of every step after the
first dropped, times offset by i * time_step_length (the Trajectory default
5.0) -- into the float32 arrays the writer stores (:196-226).

Pinned by tests/test_oracle_env.py with hand-built steps whose merged arrays
are written out by hand.
"""
from __future__ import annotations

from collections import defaultdict
from typing import Dict, List, Sequence

import numpy as np


class TrajectoryRecorder:
    def __init__(self, num_envs: int, num_units: int, saved_states: Sequence[int] = (0, 1),
                 time_step_length: float = 5.0):
        self.ne, self.n = num_envs, num_units
        self.rows = list(saved_states)
        self.T = float(time_step_length)
        self.lists: List[Dict[str, list]] = [self._empty() for _ in range(num_envs)]
        self.episodes: List[Dict[str, np.ndarray]] = []

    def _empty(self):
        return {"obs": [], "act": [], "rew": [], "time": [], "states": defaultdict(list)}

    def step(self, obs, act, rew, dense_index, dense_time, dense_state) -> None:
        """TrajectorBuffer.step (:95-127); dense_* in the solver layout (NDO, NS), (NDO, 4, NS, TILE)."""
        obs, act = np.asarray(obs), np.asarray(act)
        rew = np.asarray(rew).reshape(-1)
        for e in range(self.ne):
            L = self.lists[e]
            L["obs"].append(obs[e])
            L["act"].append(act[e])
            L["rew"].append(rew[e])
            k = int(dense_index[e])
            L["time"].append(np.array(dense_time[:k, e], dtype=np.float64))
            for r in self.rows:
                L["states"][r].append(np.array(dense_state[:k, r, e, :self.n], dtype=np.float64))
        dense_index[:] = 0

    def end_episode(self, env_ids, final_obs, lengths, returns, radii) -> None:
        """TrajectorBuffer.end_episode (:129-155) and merge_list (:38-56): one merged
        episode per env, in the order of env_ids."""
        for i, e in enumerate(np.asarray(env_ids).reshape(-1)):
            L = self.lists[int(e)]
            L["obs"].append(np.asarray(final_obs)[i])
            t = np.concatenate([d if j == 0 else d[1:] + j * self.T for j, d in enumerate(L["time"])])
            ep = {"observations": np.vstack(L["obs"]).astype(np.float32),
                  "actions": np.vstack(L["act"]).astype(np.float32),
                  "rewards": np.vstack(L["rew"]).astype(np.float32),
                  "dense_time": t.astype(np.float32),
                  "equilibrium_radii": np.asarray(radii, dtype=np.float32)[i],
                  "episode_length": int(np.asarray(lengths)[i]),
                  "episode_reward": float(np.float32(np.asarray(returns)[i])),
                  "env": int(e)}
            for r, parts in L["states"].items():
                ep[f"dense_state_{r}"] = np.concatenate(
                    [p if j == 0 else p[1:, :] for j, p in enumerate(parts)]).astype(np.float32)
            self.episodes.append(ep)
            self.lists[int(e)] = self._empty()
