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

With ``groups="env"`` every env steps on its own: the envs whose step has finished
since the last look are collected, the policy runs on them as one batch (at least
``min_batch`` envs, or whatever has waited ``max_wait_s``), and their next steps
are published at once.  A rollout then lasts about as long as the slowest env's
own sum of steps rather than the sum over steps of the slowest env.

The policy is a fixed function of an env's observation during a rollout, and a
system's arithmetic on the serving kernel does not depend on what else runs, so
every env sees the same observations, actions and failures as in the synchronous
loop (``groups=1``), whatever the number of groups and the order in which the
groups finish, as long as the policy maps each row of its batch on its own (a
GEMM's rounding may depend on the batch size).  Failed envs restart with scenes
drawn from one generator per group, so with a scene generator the results depend
on the grouping but not on the timing; with ``groups="env"`` each restart draws
from a generator of its own, seeded by (seed, env, step).
"""
from __future__ import annotations

import time
from typing import Callable, Dict, Optional

import numpy as np


class OverlappedRollout:
    def __init__(self, venv, policy: Callable[[np.ndarray, np.ndarray], np.ndarray], groups=4,
                 seed: int = 0, timeout_s: float = 300.0, min_batch: Optional[int] = None,
                 max_wait_s: float = 2e-3):
        """``policy(obs, env_ids)`` maps observations (n, N) float32 of envs ``env_ids`` to
        amplitudes (n, NK) [bar]; ``groups`` contiguous blocks of envs step independently,
        or ``"env"``: every env on its own, the policy on batches of at least ``min_batch``
        finished envs (default num_envs // 16) or on those that waited ``max_wait_s``."""
        self.venv = venv
        self.policy = policy
        self.seed = seed
        self.timeout_s = timeout_s
        self.per_env = groups == "env"
        if self.per_env:
            self.min_batch = max(1, venv.num_envs // 16) if min_batch is None else int(min_batch)
            if self.min_batch < 1:
                raise ValueError(f"min_batch must be >= 1, got {min_batch}")
            self.max_wait_s = float(max_wait_s)
            return
        if not isinstance(groups, (int, np.integer)) or groups < 1 or groups > venv.num_envs:
            raise ValueError(f"groups must be \"env\" or in [1, {venv.num_envs}], got {groups}")
        self.blocks = [b.astype(np.int32) for b in np.array_split(np.arange(venv.num_envs), groups)]
        self.rngs = [np.random.default_rng([seed, g]) for g in range(groups)]

    def collect(self, obs0: np.ndarray, steps: int) -> Dict[str, np.ndarray]:
        """``steps`` env steps of every env from observations ``obs0`` (num_envs, N).
        Returns observations (steps + 1, num_envs, N), actions (steps, num_envs, NK) [bar],
        failures (steps, num_envs), and per group the wall time at which it finished (one
        entry, the last env's, with ``groups="env"``, whose per-env times are ``env_done_s``)."""
        if self.per_env:
            return self._collect_per_env(obs0, steps)
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

    def _collect_per_env(self, obs0: np.ndarray, steps: int) -> Dict[str, np.ndarray]:
        v = self.venv
        ne, nb, nk = v.num_envs, v.num_bubbles, v.nk
        obs = np.zeros((steps + 1, ne, nb), dtype=np.float32)
        acts = np.zeros((steps, ne, nk), dtype=np.float64)
        fails = np.zeros((steps, ne), dtype=bool)
        obs[0] = obs0
        t = np.zeros(ne, dtype=np.int64)            # steps finished per env
        seq = np.zeros(ne, dtype=np.int64)          # sequence number of the env's step in flight
        flight = np.zeros(ne, dtype=bool)
        ready = np.zeros(ne, dtype=bool)            # finished, waiting for the policy
        held: list = []                             # (env ids, actions) refused by a full ring
        done_at = np.zeros(ne)
        seed = self.seed
        t0 = time.perf_counter()
        ready_since = None

        def rng_of(e):
            return np.random.default_rng([seed, e, int(t[e])])

        def act(ids):
            a = np.asarray(self.policy(obs[t[ids], ids], ids), dtype=np.float64).reshape(ids.size, nk)
            acts[t[ids], ids] = a
            held.append((ids, a))

        def publish():
            while held:
                ids, a = held[0]
                s = v.submit(ids, a)
                if s is None:
                    return
                seq[ids] = s
                flight[ids] = True
                held.pop(0)

        if steps > 0:
            act(np.arange(ne, dtype=np.int32))
            publish()
        while (t < steps).any():
            v.queue.poll()
            publish()
            moved = False
            ids = np.flatnonzero(flight).astype(np.int32)
            if ids.size:
                fin = ids[v.done_mask(ids, seq[ids])]
                if fin.size:
                    moved = True
                    flight[fin] = False
                    o, f = v.collect(fin, rng_of, record=False)
                    obs[t[fin] + 1, fin] = o
                    fails[t[fin], fin] = f
                    t[fin] += 1
                    last = fin[t[fin] >= steps]
                    done_at[last] = time.perf_counter() - t0
                    nxt = fin[t[fin] < steps]
                    if nxt.size:
                        if not ready.any():
                            ready_since = time.perf_counter()
                        ready[nxt] = True
            nready = int(ready.sum())
            if nready and (nready >= self.min_batch or not flight.any()
                           or time.perf_counter() - ready_since >= self.max_wait_s):
                r = np.flatnonzero(ready).astype(np.int32)
                ready[r] = False
                act(r)
                publish()
                moved = True
            if not moved:
                if time.perf_counter() - t0 > self.timeout_s:
                    raise TimeoutError("rollout did not finish")
                time.sleep(5e-5)
        return {"observations": obs, "actions": acts, "failures": fails,
                "group_done_s": np.array([done_at.max() if ne else 0.0]), "env_done_s": done_at}
