"""Position control of N coupled bubbles in 1-D translation: the vector environment
(reference: coupledbubble_control/envs/bubble_1D_pos_control.py:1-853).

``PosNBC1D`` has the reference's constructor, ``step`` / ``reset`` /
``reset_envs`` / ``close`` and return values (observations, rewards,
terminateds, time_outs, info), with the solver, the per-step logic and the
episode resets on the GPU:

  step(action)
    set_action            cbc_env_set_action      (one launch)
    step_physics          model.solve             (the solver's launch loop)
    observe, check_termination, get_rewards, advance_time
                          cbc_env_post            (one launch, then one sync for the counts)
    final_termination + reset_envs
                          torch.rand draws of SceneGenerator, then cbc_env_reset (one launch)

The reference runs the same sequence as some forty torch kernels with three
host synchronisations.  ``SceneGenerator`` draws the reference's uniform
stream (the same torch.Generator calls in the same order); the positions are
computed from those draws on the device.

Differences (all deliberate):
* actions are taken as float32, the dtype of the action space;
* ``reset()`` writes the new scenes into the device state every time (the
  reference writes the host buffers, which reach the device only when the
  device buffers are first created);
* the reference's default backend (Numba) is replaced by HIP; ``render_env``
  is not built;
* with ``collect_trajectories`` the episodes are recorded on the device and
  written as the reference's zarr files (envs/common/trajectory_buffer.py);
  the dense-output buffers are not zeroed after each step (only the point
  counts, which is all the solver and the recorder read).
"""
from __future__ import annotations

from typing import Dict, List, Literal, Optional, Sequence, Tuple, Union

import numpy as np
import torch

from ..backends import BackendName, KernelVariant
from ..backends.hip import _abi
from ..backends.hip.env_step import EnvKernels, ObsPart, scene_positions
from ..models import CKM1D
from .bubble_vec_env import BubbleVecEnv
from .common.specs import ActionComponent, ActionMapingManager, ObservationComponent, SpaceSchema
from .common.trajectory_buffer import TrajectorBuffer


class SceneDraws:
    """One call's scene description (cbc_env_scene) and the uniform draws it points to."""

    def __init__(self, scene: _abi.EnvSceneC, draws: Dict[str, torch.Tensor]):
        self.scene = scene
        self.draws = draws


class SceneGenerator:
    """New episode scenes (reference: bubble_1D_pos_control.py:46-249)."""

    def __init__(self, num_bubbles: int, initial_position: Union[str, List[float]],
                 target_position: Union[str, List[float]], initial_distance: Optional[float] = None,
                 target_distance: Optional[float] = None, alignment: Literal["center", "random"] = "center",
                 min_distance: float = 0.05, max_distance: float = 0.15, margin: float = 0.05,
                 domain_max: float = 1.0, domain_min: float = -1.0, seed: int = 42) -> None:
        def position(p, name):
            if isinstance(p, str):
                if p.lower() in ("random", "equidistant"):
                    return p
                raise TypeError(f"Err: {name} must be 'random', 'equidistant' (str) or a sequence of numbers.")
            if isinstance(p, list):
                if len(p) != num_bubbles:
                    raise ValueError(f"Err: {name}length is {len(p)}, but {num_bubbles} is required.")
                return [float(v) for v in p]
            raise TypeError(f"Err: {name} must be 'random' (str) or a sequence of numbers.")

        self.target_position = position(target_position, "target_position")
        self.initial_position = position(initial_position, "initial_position")
        self.initial_distance = initial_distance
        self.target_distance = target_distance
        self.alignment = alignment
        self.min_distance = min_distance
        self.max_distance = max_distance
        self.margin = margin
        self.domain_max = domain_max
        self.domain_min = domain_min
        self.num_bubbles = num_bubbles
        self._seed = seed
        self._generators: Dict[str, torch.Generator] = {}
        self.last: Optional[SceneDraws] = None

    def _get_generator(self, device) -> torch.Generator:
        device = torch.device(device)
        key = str(device)
        if key not in self._generators:
            g = torch.Generator(device=device)
            if self._seed is not None:
                g.manual_seed(self._seed)
            self._generators[key] = g
        return self._generators[key]

    @property
    def _lo_hi(self) -> Tuple[float, float]:
        return self.domain_min + self.margin, self.domain_max - self.margin

    # -- one side of a scene (cbc_scene_side) ---------------------------------
    def _random_side(self, n, dtype, device, gen, draws, key) -> _abi.SceneSideC:
        lo, hi = self._lo_hi
        available = hi - lo
        num_gaps = self.num_bubbles - 1
        if num_gaps * self.min_distance > available:
            raise ValueError(f"Cannot fit: The scene requires min length of {num_gaps * self.min_distance:.2f}, "
                             f"domain length is {available:.2f}.")
        side = _abi.SceneSideC()
        side.method = _abi.CBC_SCENE_RANDOM
        draws[key + "_gaps"] = torch.rand((n, num_gaps), dtype=dtype, device=device, generator=gen)
        if self.alignment == "random":
            side.random_align = 1
            draws[key + "_align"] = torch.rand((n, 1), dtype=dtype, device=device, generator=gen)
        elif self.alignment != "center":
            raise ValueError("alignment must be: 'center' | 'random'.")
        return side

    def _equidistant_side(self, distance, n, dtype, device, gen, draws, key) -> _abi.SceneSideC:
        lo, hi = self._lo_hi
        N = self.num_bubbles
        side = _abi.SceneSideC()
        side.method = _abi.CBC_SCENE_EQUIDISTANT
        if distance is None:
            base = torch.linspace(lo, hi, N, dtype=torch.float64)
            gap = float(base[1] - base[0])
            if gap < self.min_distance:
                raise ValueError(f"Cannot fit: The gaps between bubbles are {gap:.3f}, minimal gap is "
                                 f"{self.min_distance:.2f} ")
            if gap > self.max_distance:
                raise ValueError(f"Cannot fit: The gaps between bubbles are {gap:.3f}, maximum gap is "
                                 f"{self.max_distance:.2f} ")
            side.offset = 0.0
        elif type(distance) in (float, int):
            span = (N - 1) * distance
            if span > hi - lo:
                raise ValueError(f"Cannot fit: The scene requires min length of {span:.2f}, available domain "
                                 f"length is {hi - lo:.2f}.")
            base = torch.arange(N, dtype=torch.float64) * distance
            if self.alignment == "center":
                side.offset = (lo + hi - span) * 0.5
            elif self.alignment == "random":
                side.random_align = 1
                side.max_start = hi - span
                draws[key + "_align"] = torch.rand((n, 1), dtype=dtype, device=device, generator=gen)
            else:
                raise ValueError("alignment must be: 'lower' | 'random'.")
        else:
            raise TypeError(f"distance must be a number or None, got {type(distance)}")
        for i, v in enumerate(base.tolist()):
            side.base[i] = v
        return side

    def _fixed_side(self, position) -> _abi.SceneSideC:
        side = _abi.SceneSideC()
        side.method = _abi.CBC_SCENE_FIXED
        for i, v in enumerate(torch.as_tensor(position, dtype=torch.float64).tolist()):
            side.base[i] = v
        return side

    def draw(self, num_scenes: int, dtype: torch.dtype = torch.float64, device="cpu",
             kernel_device=None) -> SceneDraws:
        """Draws of one generate_new_scenes call (:112-137) with the generator of
        ``device``; the scene description points at copies on ``kernel_device``."""
        gen = self._get_generator(device)
        draws: Dict[str, torch.Tensor] = {}
        n = int(num_scenes)
        if self.initial_position == "random":
            b = self._random_side(n, dtype, device, gen, draws, "bubbles")
        elif self.initial_position == "equidistant":
            b = self._equidistant_side(self.initial_distance, n, dtype, device, gen, draws, "bubbles")
        else:
            b = self._fixed_side(self.initial_position)
            self._random_side(n, dtype, device, gen, {}, "burn")    # the reference draws and discards
        if self.target_position == "random":
            t = self._random_side(n, dtype, device, gen, draws, "targets")
        elif self.target_position == "equidistant":
            # the reference spaces targets by initial_distance (:131)
            t = self._equidistant_side(self.initial_distance, n, dtype, device, gen, draws, "targets")
        else:
            t = self._fixed_side(self.target_position)
        kdev = torch.device(kernel_device) if kernel_device is not None else torch.device(device)
        dev_draws = {k: v.to(device=kdev, dtype=torch.float64).contiguous() for k, v in draws.items()}
        lo, hi = self._lo_hi
        sc = _abi.EnvSceneC()
        sc.lo = lo
        sc.available = hi - lo
        sc.min_distance = self.min_distance
        sc.room_per_gap = self.max_distance - self.min_distance
        sc.extra_gap_limit = (hi - lo) - (self.num_bubbles - 1) * self.min_distance
        for side, key in ((b, "bubbles"), (t, "targets")):
            if key + "_gaps" in dev_draws:
                side.u_gaps = _abi.ptr(dev_draws[key + "_gaps"])
            if key + "_align" in dev_draws:
                side.u_align = _abi.ptr(dev_draws[key + "_align"])
        sc.bubbles, sc.targets = b, t
        self.last = SceneDraws(sc, dev_draws)
        return self.last

    def generate_new_scenes(self, num_scenes, dtype: torch.dtype = torch.float32,
                            device: str = "cpu") -> Tuple[torch.Tensor, torch.Tensor]:
        """(bubble_positions, target_positions), each (num_scenes, N), on ``device``."""
        kdev = torch.device("cuda") if torch.device(device).type == "cpu" else torch.device(device)
        d = self.draw(num_scenes, dtype, device, kdev)
        pos, tgt = scene_positions(d.scene, int(num_scenes), self.num_bubbles, kdev)
        return pos.to(device=device, dtype=dtype), tgt.to(device=device, dtype=dtype)


def _default_action_schema() -> SpaceSchema:
    s = SpaceSchema()
    s.components = [ActionComponent(name="PA", idx=[0, 1], min=[0.0, 0.0], max=[1.0, 1.0], scale=1.0e5),
                    ActionComponent(name="PS", idx=[0, 1], min=[0.0, 0.0], max=[0.25, 0.25], scale=np.pi)]
    return s


def _default_observation_schema(n: int) -> SpaceSchema:
    s = SpaceSchema()
    s.components = [ObservationComponent(name="XT", idx=list(range(n)), min=[-0.5] * n, max=[0.5] * n),
                    ObservationComponent(name="X", idx=list(range(n)), min=[-0.5] * n, max=[0.5] * n, stack=2)]
    return s


class PosNBC1D(BubbleVecEnv):
    """Pos: position control, NB: N bubbles, C: coupled, 1D: one-dimensional translation."""

    def __init__(
            self,
            num_bubbles: int,
            num_envs: int,
            envs_per_block: int,
            action_space_schema: Optional[SpaceSchema] = None,
            observation_space_schema: Optional[SpaceSchema] = None,
            R0: Optional[List[float]] = None,
            PA: Optional[List[float]] = None,
            FR: Optional[List[float]] = None,
            PS: Optional[List[float]] = None,
            REF_FREQ: float = 25.0,
            num_components: int = 2,
            acoustic_field: Literal["CONST", "SW_N", "SW_A"] = "SW_N",
            episode_length: int = 128,
            time_step_length: Union[int, float] = 5.0,
            seed: int = 42,
            target_position: Union[str, List[float]] = "random",
            initial_position: Union[str, List[float]] = "random",
            initial_distance: Optional[float] = None,
            target_distance: Optional[float] = None,
            alignment: Literal["center", "random"] = "center",
            min_distance: float = 0.05,
            max_distance: float = 0.9,
            margin: float = 0.05,
            dnorm: float = 0.2,
            position_tolerance: float = 0.01,
            apply_termination: bool = True,
            rewards_weights: Optional[List[float]] = None,
            reward_exp: float = 0.5,
            positive_terminal_reward: float = 10.0,
            negative_terminal_reward: float = -10.0,
            render_env: bool = False,
            collect_trajectories: bool = True,
            trajectory_resolution: int = 128,
            trajectory_buffer_store: str = "PosNBC1D",
            backend: BackendName = BackendName.HIP,
            variant: KernelVariant = KernelVariant.WARP,
            max_kernel_steps: int = int(1e6),
            kernel_steps: int = 2048,
            compiler: str = "hipcc",
            max_registers: Optional[int] = None,
            fastmath: bool = True,
            cuda_mode: Literal["debug", "profile", "release"] = "release",
            device: str = "cuda",
            keep_step: bool = False,
            verbose: bool = True,
            trajectory_async_write: bool = False,
    ) -> None:
        """Arguments as the reference's (:260-354).  Extra: ``keep_step`` keeps each
        env's last accepted step size across steps instead of resetting it to 1e-6;
        ``verbose`` prints the per-step failure / success / time-out counts as the
        reference does; ``trajectory_async_write`` writes episode files on a
        background thread (``close()`` waits for it)."""
        if render_env:
            raise NotImplementedError("render_env: rendering is not part of this package")
        self.action_space_schema = action_space_schema if action_space_schema is not None else _default_action_schema()
        self.observation_space_schema = (observation_space_schema if observation_space_schema is not None
                                         else _default_observation_schema(num_bubbles))
        super().__init__(num_envs, envs_per_block, self.action_space_schema, self.observation_space_schema, seed,
                         device)

        def init_list(x, size, name, default, scale: float = 1.0):
            src = x if x is not None else default
            if len(src) > size:
                raise ValueError(f"{name}: expected length {size}, got {len(src)}")
            return [float(v) * scale for v in src]

        self.default_R0 = init_list(R0, num_bubbles, "R0", [60.0] * num_bubbles)
        self.default_PA = init_list(PA, num_components, "PA", [1.0, 0.0])
        self.default_FR = init_list(FR, num_components, "FR", [25.0, 50.0])
        self.default_PS = init_list(PS, num_components, "PS", [0.0, 0.0])
        self.REF_FREQ = REF_FREQ
        ne = self.num_envs

        def tile_col(v):
            return np.tile(np.asarray(v, dtype=np.float64)[:, np.newaxis], (1, ne))

        self.model = CKM1D(
            num_systems=ne, systems_per_block=self.envs_per_block, units_per_system=num_bubbles,
            num_components=num_components, acoustic_field=acoustic_field,
            num_stored_points=trajectory_resolution if collect_trajectories else 0,
            R0=np.tile(np.asarray(self.default_R0, dtype=np.float64), (ne, 1)),
            PA=tile_col(self.default_PA), FR=tile_col(self.default_FR), PS=tile_col(self.default_PS),
            REF_FREQ=self.REF_FREQ, backend=backend, variant=variant, cuda_mode=cuda_mode, compiler=compiler,
            max_registers=max_registers, fastmath=fastmath)
        self.solver_kwargs = {"debug": False, "max_steps": max_kernel_steps, "kernel_steps": kernel_steps}
        self._action_map = ActionMapingManager(self.action_space_schema, self.model.control_layout)

        self.episode_length = episode_length
        self.time_step_length = time_step_length
        self.seed = seed
        self.num_components = num_components
        self.num_bubbles = num_bubbles
        self.device = device
        self.verbose = verbose

        comps = self.observation_space_schema.components
        self.domain_max = max(max(c.max) for c in comps if c.name in ("X", "XT"))
        self.domain_min = min(min(c.min) for c in comps if c.name in ("X", "XT"))
        self.scene_generator = SceneGenerator(self.num_bubbles, initial_position, target_position, initial_distance,
                                              target_distance, alignment, min_distance, max_distance, margin,
                                              self.domain_max, self.domain_min, seed=self.seed)
        self.reward_params = {
            "weights": torch.tensor(init_list(rewards_weights, 2, "reward_weights", [1.0, 0.0]),
                                    dtype=torch.float32, device=device),
            "reward_exp": reward_exp,
            "r_dnorm": 1.0 / dnorm,
            "positive_terminal_reward": positive_terminal_reward,
            "negative_terminal_reward": negative_terminal_reward,
        }
        self.apply_termination = apply_termination
        self.position_tolerance = position_tolerance
        self._collect_trajectories = collect_trajectories
        self.trajectory_buffer = TrajectorBuffer(
            num_envs=self.num_envs, num_units=self.num_bubbles, saved_states=[0, 1],
            zarr_root=trajectory_buffer_store, max_episode_steps=episode_length, verbose=verbose,
            async_write=trajectory_async_write,
            metadata={
                "num_bubbles": self.num_bubbles,
                "ac_type": acoustic_field,
                "k": num_components,
                "PA_max": self.action_space_schema.components[0].max,
                "PA_min": self.action_space_schema.components[0].min,
                "PS_max": self.action_space_schema.components[1].max,
                "PS_min": self.action_space_schema.components[1].max,    # as the reference stores it
                "freq": self.default_FR,
                "ref_freq": self.REF_FREQ,
                "XT_max": self.observation_space_schema.components[0].max,
                "XT_min": self.observation_space_schema.components[0].min,
                "X_max": self.observation_space_schema.components[1].max,
                "X_min": self.observation_space_schema.components[1].max,    # as the reference stores it
                "position_tolerance": self.position_tolerance,
                "time_step_length": self.time_step_length,
            }) if collect_trajectories else None

        parts = []
        for c in comps:
            if len(c.idx) != num_bubbles:
                raise ValueError(f"observation component {c.name!r} must cover all {num_bubbles} bubbles")
            if c.name == "XT":
                if c.stack != 1:
                    raise ValueError("the target component XT cannot be stacked")
                parts.append(ObsPart(c.name, _abi.CBC_ENV_PART_TARGET, 0, 1))
            elif c.name in self.model.state_layout:
                parts.append(ObsPart(c.name, _abi.CBC_ENV_PART_STATE, self.model.state_layout[c.name].base_offset,
                                     int(c.stack)))
            else:
                parts.append(ObsPart(c.name, _abi.CBC_ENV_PART_FIXED, 0, int(c.stack)))
        names = [p.name for p in parts]
        if "X" not in names or "XT" not in names:
            raise ValueError("the observation schema needs the components 'X' and 'XT'")
        w = self.reward_params["weights"].cpu().tolist()
        self._kernels = EnvKernels(
            self.model, parts, self.observation_space._bias, self.observation_space._rscale,
            action_rows=self._action_map._dev_cols, action_scales=self._action_map._scales,
            pa_cols=list(self._action_map._feat_col.get("PA", [])), position_tolerance=position_tolerance,
            dnorm=dnorm, reward_exp=reward_exp, weights=w, positive_terminal_reward=positive_terminal_reward,
            negative_terminal_reward=negative_terminal_reward, apply_termination=apply_termination,
            episode_length=episode_length, next_time_step=0.0 if keep_step else 1e-6)
        self._observed = 0
        self._observations: Optional[torch.Tensor] = None
        self._actions: Optional[torch.Tensor] = None
        self._build_interop_views()

    # -- properties (:621-636) -------------------------------------------------
    @property
    def algo_steps(self) -> torch.Tensor:
        return self._kernels.algo_steps

    @property
    def total_rewards(self) -> torch.Tensor:
        return self._kernels.total_rewards

    @property
    def target_positions(self) -> torch.Tensor:
        return self._kernels.part_view("XT")

    @property
    def bubble_positions(self) -> torch.Tensor:
        ring = self._kernels.part_view("X")
        return ring[(self._observed - 1) % ring.shape[0]]

    @property
    def target_distance(self) -> torch.Tensor:
        return torch.abs(self.target_positions - self.bubble_positions)

    def _build_interop_views(self):
        """The solver buffers the step touches (:708-723); torch tensors here."""
        db = self.model.runtime.device_buffers
        self._interop = {"time_step": db.state.time_step, "actual_time": db.state.actual_time,
                         "actual_state": db.state.actual_state, "control_params": db.params.control_params,
                         "status_flags": db.status.status_flags, "actual_event": db.status.actual_event}

    # -- the step (:503-538) ---------------------------------------------------
    @torch.no_grad()
    def step(self, action: Optional[torch.Tensor] = None, clone: bool = False):
        if self._observations is None:
            raise RuntimeError("call reset() before step()")
        if action is None:
            action = self.action_space.sample()
        self.set_action(action)
        self.step_physics()
        self._post_step()
        if self.trajectory_buffer is not None:
            dd = self.model.runtime.device_buffers.dense
            self.trajectory_buffer.step(self._observations, self._actions, self._rewards, dd.dense_index,
                                        dd.dense_time, dd.dense_state)
        info = self.final_termination()
        out = (self._observations, self._rewards, self._terminateds, self._time_outs, info)
        if clone:
            return tuple(t.clone() for t in out[:4]) + (info,)
        return out

    @torch.no_grad()
    def set_action(self, actions: torch.Tensor, ptr: Optional[int] = None, size: Optional[int] = None):
        self._actions = actions
        self._kernels.set_action(actions)

    def step_physics(self):
        self.model.solve(**self.solver_kwargs)

    def _post_step(self):
        """observe, check_termination, get_rewards, advance_time (:756-827)."""
        c = self._kernels.post(self._observed)
        self._observed += 1
        o = self._kernels.outputs
        self._observations, self._rewards = o.observations, o.rewards
        self._failure, self._success = o.failure, o.success
        self._terminateds, self._time_outs = o.terminated, o.time_outs
        self._counts = c
        if self.verbose:
            if c.failures:
                print("Number of episode failures:", c.failures)
            if c.successes:
                print("Number of episode success:", c.successes)
            if c.time_outs:
                print("Number of episode time-outs:", c.time_outs)

    @torch.no_grad()
    def final_termination(self):
        """Done envs' final outputs, then their reset (:830-853)."""
        k = int(self._counts.done)
        if k == 0:
            return {}
        ids = self._kernels.done_ids[:k].clone()
        d = self.scene_generator.draw(k, torch.float64, self.device, self._kernels.device)
        fo, fr, fl = self._kernels.reset(d.scene, ids, self._observed, final=True)
        if self.trajectory_buffer is not None:
            self.trajectory_buffer.end_episode(ids, fo, fl, fr, np.tile(self.default_R0, (k, 1)))
        return {"final_observation": fo, "dones": ids, "episode_return": fr, "episode_length": fl}

    # -- resets (:541-614) -----------------------------------------------------
    @torch.no_grad()
    def reset(self, **kwargs):
        """New scenes for every env from the CPU generator; counters cleared."""
        hb = self.model.runtime.host_buffers
        hb.state.time_end[:] = self.time_step_length
        self.model._ensure_buffers()
        db = self.model.runtime.device_buffers
        db.state.time_end.fill_(float(self.time_step_length))
        d = self.scene_generator.draw(self.num_envs, torch.float64, "cpu", self._kernels.device)
        ids = torch.arange(self.num_envs, dtype=torch.int64, device=self._kernels.device)
        obs = torch.empty((self.num_envs, self._kernels.obs_dim), dtype=torch.float32, device=self._kernels.device)
        self._kernels.reset(d.scene, ids, self._observed, final=False, observations=obs)
        self._observations = obs
        hb.state.actual_state[:] = db.state.actual_state.cpu().numpy()
        self._build_interop_views()
        return self._observations.clone(), {}

    @torch.no_grad()
    def reset_envs(self, env_ids, sync_host: bool = False, **kwargs):
        """New scenes for ``env_ids`` from the device generator (:583-614)."""
        ids = torch.as_tensor(env_ids, dtype=torch.int64, device=self._kernels.device).reshape(-1)
        d = self.scene_generator.draw(int(ids.numel()), torch.float64, self.device, self._kernels.device)
        obs = self._observations.clone()
        self._kernels.reset(d.scene, ids, self._observed, final=False, observations=obs)
        self._observations = obs
        if sync_host:
            idx = ids.cpu().numpy()
            hb = self.model.runtime.host_buffers
            x = self.model.runtime.device_buffers.state.actual_state[1].cpu().numpy()
            hb.state.actual_state[1, idx, :self.num_bubbles] = x[idx, :self.num_bubbles]

    @torch.no_grad()
    def close(self):
        if self.trajectory_buffer is not None:
            self.trajectory_buffer.close()
        print("close")

    def render(self):
        raise NotImplementedError("rendering is not part of this package")
