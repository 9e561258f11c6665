"""Rollout collection with the policy overlapped with solver serving.

The reference's PPO collects a rollout (rl/ppo.py:249-300) by alternating, for one
synchronous vector env, the policy forward on every env's observation and
``env.step`` on every env: the GPU is idle during inference and every env waits
for the slowest system of the whole batch at every step.  ``OverlappedRollout``
splits the envs of a ``ServedVecEnv`` into groups that step on their own: when a
group's step has finished, the policy runs on that group's observations and the
group's next step is published at once, while the other groups' steps keep
running on the serving kernel.  Inference overlaps the solver, and a group waits
only for its own slowest system.

The policy is a fixed function of an env's observation during a rollout, and a
system's arithmetic on the serving kernel does not depend on what else runs, so
every env sees the same observations, actions and failures as in the synchronous
loop (``groups=1``), whatever the number of groups and the order in which the
groups finish.  Failed envs restart with scenes drawn from one generator per
group, so with a scene generator the results depend on the grouping but not on
the timing.
"""
from __future__ import annotations

import time
from typing import Callable, Dict, Optional

import numpy as np


class OverlappedRollout:
    def __init__(self, venv, policy: Callable[[np.ndarray, np.ndarray], np.ndarray], groups: int = 4,
                 seed: int = 0, timeout_s: float = 300.0):
        """``policy(obs, env_ids)`` maps observations (n, N) float32 of envs ``env_ids`` to
        amplitudes (n, NK) [bar]; ``groups`` contiguous blocks of envs step independently."""
        if groups < 1 or groups > venv.num_envs:
            raise ValueError(f"groups must be in [1, {venv.num_envs}], got {groups}")
        self.venv = venv
        self.policy = policy
        self.blocks = [b.astype(np.int32) for b in np.array_split(np.arange(venv.num_envs), groups)]
        self.rngs = [np.random.default_rng([seed, g]) for g in range(groups)]
        self.timeout_s = timeout_s

    def collect(self, obs0: np.ndarray, steps: int) -> Dict[str, np.ndarray]:
        """``steps`` env steps of every env from observations ``obs0`` (num_envs, N).
        Returns observations (steps + 1, num_envs, N), actions (steps, num_envs, NK) [bar],
        failures (steps, num_envs), and per group the wall time at which it finished."""
        v = self.venv
        ne, nb, nk = v.num_envs, v.num_bubbles, v.nk
        obs = np.zeros((steps + 1, ne, nb), dtype=np.float32)
        acts = np.zeros((steps, ne, nk), dtype=np.float64)
        fails = np.zeros((steps, ne), dtype=bool)
        obs[0] = obs0
        G = len(self.blocks)
        t = [0] * G                     # steps finished per group
        seq: list = [None] * G          # sequence numbers of the group's step in flight
        want: list = [None] * G         # actions waiting for ring slots
        done_at = np.zeros(G)
        t0 = time.perf_counter()

        def act(g):
            blk = self.blocks[g]
            a = np.asarray(self.policy(obs[t[g], blk], blk), dtype=np.float64).reshape(blk.size, nk)
            acts[t[g], blk] = a
            want[g] = a

        def publish(g):
            s = v.submit(self.blocks[g], want[g])
            if s is not None:
                seq[g], want[g] = s, None

        if steps > 0:
            for g in range(G):
                act(g)
                publish(g)
        while min(t) < steps:
            v.queue.poll()
            moved = False
            for g in range(G):
                if t[g] >= steps:
                    continue
                if want[g] is not None:          # ring was full: publish as slots free up
                    publish(g)
                    moved = moved or want[g] is None
                    continue
                blk = self.blocks[g]
                if not v.finished(blk, seq[g]):
                    continue
                o, f = v.collect(blk, self.rngs[g])
                obs[t[g] + 1, blk] = o
                fails[t[g], blk] = f
                t[g] += 1
                moved = True
                if t[g] < steps:
                    act(g)
                    publish(g)
                else:
                    done_at[g] = time.perf_counter() - t0
            if not moved:
                if time.perf_counter() - t0 > self.timeout_s:
                    raise TimeoutError("rollout did not finish")
                time.sleep(5e-5)
        return {"observations": obs, "actions": acts, "failures": fails, "group_done_s": done_at}
