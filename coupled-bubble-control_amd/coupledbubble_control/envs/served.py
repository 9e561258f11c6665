"""The solver side of the vector environment's step over a serving queue.

The reference's ``PosNBC1D.step`` (envs/bubble_1D_pos_control.py:503-538) runs, on
the model's device buffers, ``set_action`` (amplitudes into control_params,
:731-749), ``step_physics`` (a blocking ``model.solve``, :751-753), ``observe``
(the X rows, :755-769), ``check_termination`` (failure = status != SUCCESS,
:771-789), ``advance_time`` (time 0, step size 1e-6, event -1, :815-826) and, for
terminated envs, ``reset_envs`` (new positions, R/R0 = 1, U = V = 0, radii kept,
:583-614).  ``ServedVecEnv.step`` does the same through a ``ServeQueue``: one
item per env carrying its amplitudes and the step length, the step size reset
to 1e-6 as ``advance_time`` does (or kept with ``keep_step``), terminated envs
restarted through the staging mirrors on their next item.  Observations come
from the host-mapped end states, so a step needs no device synchronisation.
Reward shaping, targets and observation stacking stay with the environment.
"""
from __future__ import annotations

from typing import Callable, Dict, Optional, Tuple

import numpy as np


class ServedVecEnv:
    def __init__(self, model, time_step_length: float = 5.0, keep_step: bool = False,
                 scene: Optional[Callable[[int, np.random.Generator], np.ndarray]] = None, seed: int = 0,
                 opts=None, timeout_s: float = 120.0):
        """``scene(k, rng)`` returns new positions (k, N) [wavelengths] for reset envs
        (the scene generator); None keeps terminated envs running."""
        from ..backends.hip.serve import ServeOptions, ServeQueue

        self.model = model
        self.num_envs = model._execution_spec.ns
        self.num_bubbles = model._execution_spec.ups
        self.nk = model._model_spec.nk
        self.T = float(time_step_length)
        self.keep_step = keep_step
        self.scene = scene
        self.rng = np.random.default_rng(seed)
        self.timeout_s = timeout_s
        cap = 64
        while cap < 2 * self.num_envs:
            cap *= 2
        o = opts or ServeOptions()
        self.queue = ServeQueue(model, ServeOptions(**{**o.__dict__, "capacity": max(o.capacity, cap)}))
        self._ids = np.arange(self.num_envs, dtype=np.int32)
        self._reset = np.zeros(self.num_envs, dtype=bool)
        self.observations: Optional[np.ndarray] = None
        self.failures: Optional[np.ndarray] = None

    def step(self, pa_bar: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
        """One step of every env with amplitudes ``pa_bar`` (num_envs, NK) [bar]; returns the
        observed positions (num_envs, N) as float32 and the failure mask."""
        import time

        seq = self.submit(self._ids, pa_bar)
        if seq is None:
            raise RuntimeError("serving ring full")
        t0 = time.time()
        while not self.finished(self._ids, seq):
            self.queue.poll()
            if time.time() - t0 > self.timeout_s:
                raise TimeoutError("env step did not finish")
            time.sleep(1e-4)
        self.collect(self._ids)
        return self.observations, self.failures

    def submit(self, env_ids: np.ndarray, pa_bar: np.ndarray) -> Optional[np.ndarray]:
        """Publish one step of ``env_ids`` with amplitudes ``pa_bar`` (len(env_ids), NK) [bar]
        without waiting; returns the items' sequence numbers, or None when the ring has too
        few free slots (nothing published: poll and try again)."""
        ids = np.ascontiguousarray(env_ids, dtype=np.int32)
        pa = np.ascontiguousarray(np.asarray(pa_bar, dtype=np.float64).reshape(ids.size, self.nk).T) * 1.0e5
        rs = self._reset[ids]
        h0 = np.where(rs, 1e-6, 0.0 if self.keep_step else 1e-6)
        seq = self.queue.submit(ids, pa, self.T, h0, rs if rs.any() else None)
        if seq is not None:
            self._reset[ids] = False
        return seq

    def finished(self, env_ids: np.ndarray, seq: np.ndarray) -> bool:
        """Whether the step with sequence numbers ``seq`` of ``env_ids`` has finished."""
        return bool(self.queue.done(env_ids, seq).all())

    def done_mask(self, env_ids: np.ndarray, seq: np.ndarray) -> np.ndarray:
        """Per env of ``env_ids``: whether its step with sequence number ``seq`` has finished."""
        return self.queue.done(env_ids, seq)

    def collect(self, env_ids: np.ndarray, rng=None, record: bool = True) -> Tuple[np.ndarray, np.ndarray]:
        """Observations (len(env_ids), N) and failures of the finished step of ``env_ids``
        (also written into ``observations`` / ``failures`` unless ``record`` is False);
        failed envs get new scenes for their next step, drawn from ``rng``: a generator,
        a function of the env id returning one (one draw per env), or None (the env's
        own generator)."""
        ids = np.asarray(env_ids)
        state, misc = self.queue.results(ids)
        obs = state[1].astype(np.float32)
        fail = misc[:, 2] != 1.0
        if record:
            # new arrays each time: observations returned by earlier steps stay as they were
            prev_o = self.observations if self.observations is not None else \
                np.zeros((self.num_envs, self.num_bubbles), dtype=np.float32)
            prev_f = self.failures if self.failures is not None else np.zeros(self.num_envs, dtype=bool)
            self.observations, self.failures = prev_o.copy(), prev_f.copy()
            self.observations[ids] = obs
            self.failures[ids] = fail
        if self.scene is not None and fail.any():
            self.reset_envs(ids[fail], rng)
        return obs, fail

    def reset_envs(self, env_ids, rng=None) -> None:
        """New positions for ``env_ids`` (radii kept), taken by their next step; ``rng`` as
        for ``collect``."""
        env_ids = np.asarray(env_ids, dtype=np.int64)
        if callable(rng):
            X0 = np.concatenate([self.scene(1, rng(int(e))) for e in env_ids], axis=0)
        else:
            X0 = self.scene(env_ids.size, self.rng if rng is None else rng)
        state, up, cm = self.model.scene_rows(env_ids, None, X0)
        self.queue.stage(env_ids, state, up, cm)
        self._reset[env_ids] = True

    def close(self) -> None:
        self.queue.close()


class ServedPosNBC1D:
    """``PosNBC1D.step`` (envs/bubble_1D_pos_control.py:503-538) served group by group.

    The envs of a ``PosNBC1D`` are split into ``groups`` contiguous blocks that step on
    their own.  A block's step is the reference step with the solver call replaced by
    serving items: ``set_action`` on the block's rows (cbc_env_set_action with an env
    range: amplitudes and phases into control_params), one item per env that takes its
    amplitudes from control_params and integrates ``time_step_length`` from the resident
    state with the step size ``advance_time`` left, then -- once the block's items have
    finished -- ``observe``, ``check_termination``, ``get_rewards`` and ``advance_time`` on
    the block's rows (cbc_env_post with an env range) and ``final_termination`` with
    ``reset_envs`` for its done envs (cbc_env_reset, scenes from the block's own scene
    generator).  While one block is being observed, reset and given its next actions,
    the other blocks' items keep integrating on the serving kernel.

    A system's arithmetic on the serving kernel equals the plain launch's, so with one
    block a rollout equals ``PosNBC1D.step`` in a loop bit for bit (the block's scene
    generator is the environment's, seeded alike); with several it equals stepping the
    blocks in lockstep, whatever order they finish in.  Block g > 0 draws its scenes
    from a generator seeded ``seed + 1000 g``.  Trajectory recording is not served
    (the serving kernels record no dense output)."""

    def __init__(self, env, groups: int = 1, opts=None, timeout_s: float = 300.0):
        import copy

        import torch

        from ..backends.hip.serve import ServeOptions, ServeQueue

        if env.trajectory_buffer is not None:
            raise ValueError("served steps record no trajectories: build the env with collect_trajectories=False")
        if env._observations is None:
            raise RuntimeError("call env.reset() first")
        ne = env.num_envs
        if not isinstance(groups, (int, np.integer)) or not 1 <= groups <= ne:
            raise ValueError(f"groups must be in [1, {ne}], got {groups}")
        self.env = env
        self.kernels = env._kernels
        self.num_envs = ne
        self.T = float(env.time_step_length)
        self.timeout_s = float(timeout_s)
        self.blocks = [(int(b[0]), int(b.size)) for b in np.array_split(np.arange(ne), groups)]
        self.ids = [np.arange(b, b + c, dtype=np.int32) for b, c in self.blocks]
        self.scenes = []
        for g in range(groups):
            sg = copy.copy(env.scene_generator)
            sg._generators = {}
            sg._seed = env.scene_generator._seed + 1000 * g
            sg.last = None
            self.scenes.append(sg)
        self.observed = [env._observed] * groups
        cap = 64
        while cap < 2 * ne:
            cap *= 2
        o = opts or ServeOptions()
        self.queue = ServeQueue(env.model, ServeOptions(**{**o.__dict__, "capacity": max(o.capacity, cap)}))
        self.out = self.kernels.new_outputs()
        self.out.observations.copy_(env._observations)
        self.torch = torch
        self.draw_log: Optional[list] = None     # (group, step, done ids, draws) when recording

    # -- one block ---------------------------------------------------------------
    def _start(self, g: int, actions):
        b, c = self.blocks[g]
        self.kernels.set_action_range(b, actions)
        self.torch.cuda.current_stream().synchronize()     # control rows written before the items
        return self.queue.submit(self.ids[g], None, self.T, 0.0, use_control=True)

    def _finish(self, g: int, step: int):
        """observe .. reset of block g's finished step; returns (counts, done ids, final outputs)."""
        b, c = self.blocks[g]
        counts = self.kernels.post_range(b, c, self.observed[g], self.out)
        self.observed[g] += 1
        k = int(counts.done)
        if k == 0:
            return counts, None, None
        ids = self.kernels.done_ids[:k].clone()
        d = self.scenes[g].draw(k, self.torch.float64, self.env.device, self.kernels.device)
        if self.draw_log is not None:
            self.draw_log.append((g, step, ids.cpu().numpy(), {n: v.cpu().numpy() for n, v in d.draws.items()}))
        fin = self.kernels.reset(d.scene, ids, self.observed[g], final=True, observations=self.out.observations)
        return counts, ids, fin

    # -- rollouts -----------------------------------------------------------------
    def rollout(self, policy, steps: int, lockstep: bool = False, record: bool = True) -> Dict:
        """``steps`` env steps of every env.  ``policy(obs, begin)`` maps a block's
        observation rows (n, obs_dim) float32 on the device to its actions (n, A).  Blocks
        step on their own (``lockstep=False``) or all together (the synchronous loop of
        rl/ppo.py:249-300: every block steps, then the policy runs on every block).
        Returns observations (steps + 1, NE, obs_dim), actions (steps, NE, A), rewards,
        terminated, time_outs (steps, NE) as CPU tensors when ``record``, the done ids
        per (step, block) and the seconds at which each block finished."""
        import time

        torch = self.torch
        ne, G = self.num_envs, len(self.blocks)
        obs_dim = self.kernels.obs_dim
        A = self.kernels.spec.num_actions
        rec = None
        if record:
            rec = {"observations": torch.zeros((steps + 1, ne, obs_dim)),
                   "actions": torch.zeros((steps, ne, A)), "rewards": torch.zeros((steps, ne)),
                   "terminated": torch.zeros((steps, ne), dtype=torch.bool),
                   "time_outs": torch.zeros((steps, ne), dtype=torch.bool)}
            rec["observations"][0] = self.out.observations.cpu()
        dones: Dict = {}
        t = [0] * G
        seq = [None] * G
        pending = [None] * G             # actions waiting for ring slots
        done_at = np.zeros(G)
        t0 = time.perf_counter()

        def act_and_submit(g):
            b, c = self.blocks[g]
            if pending[g] is None:
                a = policy(self.out.observations[b:b + c], b)
                if rec is not None:
                    rec["actions"][t[g], b:b + c] = a.detach().float().cpu()
                pending[g] = a
            s = self._start(g, pending[g])
            if s is not None:
                seq[g] = s
                pending[g] = None

        def finish(g):
            b, c = self.blocks[g]
            counts, ids, fin = self._finish(g, t[g])
            if rec is not None:
                o = self.out
                rec["rewards"][t[g], b:b + c] = o.rewards[b:b + c].cpu()
                rec["terminated"][t[g], b:b + c] = o.terminated[b:b + c].cpu()
                rec["time_outs"][t[g], b:b + c] = o.time_outs[b:b + c].cpu()
                rec["observations"][t[g] + 1, b:b + c] = o.observations[b:b + c].cpu()
            if ids is not None:
                dones[(t[g], g)] = (ids.cpu(), [x.cpu() for x in fin])
            seq[g] = None
            t[g] += 1
            if t[g] == steps:
                done_at[g] = time.perf_counter() - t0

        if steps > 0:
            for g in range(G):
                act_and_submit(g)
        while min(t) < steps:
            if time.perf_counter() - t0 > self.timeout_s:
                raise TimeoutError(f"rollout did not finish in {self.timeout_s} s (steps done {t})")
            progressed = False
            if lockstep:
                if all(seq[g] is not None and self.queue.done(self.ids[g], seq[g]).all() for g in range(G)):
                    for g in range(G):
                        finish(g)
                    if min(t) < steps:
                        for g in range(G):
                            act_and_submit(g)
                    progressed = True
                elif any(seq[g] is None for g in range(G)):
                    for g in range(G):
                        if seq[g] is None:
                            act_and_submit(g)
            else:
                for g in range(G):
                    if t[g] >= steps:
                        continue
                    if seq[g] is None:
                        act_and_submit(g)
                    elif self.queue.done(self.ids[g], seq[g]).all():
                        finish(g)
                        progressed = True
                        if t[g] < steps:
                            act_and_submit(g)
            if not progressed:
                self.queue.poll()
                time.sleep(5e-5)
        out = {"dones": dones, "group_done_s": done_at, "wall_s": time.perf_counter() - t0}
        if rec is not None:
            out.update(rec)
        return out

    def close(self) -> None:
        self.queue.close()
