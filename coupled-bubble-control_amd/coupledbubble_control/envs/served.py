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

from typing import Callable, Optional, Tuple

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
