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

        pa = np.ascontiguousarray(np.asarray(pa_bar, dtype=np.float64).reshape(self.num_envs, self.nk).T) * 1.0e5
        h0 = np.where(self._reset, 1e-6, 0.0 if self.keep_step else 1e-6)
        seq = self.queue.submit(self._ids, pa, self.T, h0, self._reset if self._reset.any() else None)
        if seq is None:
            raise RuntimeError("serving ring full")
        self._reset[:] = False
        t0 = time.time()
        while not self.queue.done(self._ids, seq).all():
            self.queue.poll()
            if time.time() - t0 > self.timeout_s:
                raise TimeoutError("env step did not finish")
            time.sleep(1e-4)
        state, misc = self.queue.results(self._ids)
        self.observations = state[1].astype(np.float32)
        self.failures = misc[:, 2] != 1.0
        if self.scene is not None and self.failures.any():
            self.reset_envs(np.nonzero(self.failures)[0])
        return self.observations, self.failures

    def reset_envs(self, env_ids) -> None:
        """New positions for ``env_ids`` (radii kept), taken by their next step."""
        env_ids = np.asarray(env_ids, dtype=np.int64)
        X0 = self.scene(env_ids.size, self.rng)
        state, up, cm = self.model.scene_rows(env_ids, None, X0)
        self.queue.stage(env_ids, state, up, cm)
        self._reset[env_ids] = True

    def close(self) -> None:
        self.queue.close()
