"""ctypes binding of libcbc.so (C ABI declared in include/cbc.h).

The library is loaded from ``coupled-bubble-control_amd/lib/libcbc.so``.
``import torch`` happens first so that the process has a single HIP runtime
(torch's); the library's DT_NEEDED ``libamdhip64.so.7`` then resolves to it.
There is no fallback: if the library is missing, every call raises.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import torch  # noqa: F401  (load torch's HIP runtime before libcbc)

_PKG_ROOT = Path(__file__).resolve().parents[3]
LIB_PATH = Path(os.environ.get("CBC_LIB", _PKG_ROOT / "lib" / "libcbc.so"))

CBC_ABI_VERSION = 5

CBC_OK = 0
CBC_ERR_INVALID = -1
CBC_ERR_HIP = -2
CBC_ERR_UNSUPPORTED = -3
CBC_ERR_NOMEM = -4
CBC_ERR_BUSY = -5

CBC_ACTIVE = 0
CBC_SUCCESS = 1
CBC_EVENT_TERMINAL = 2
CBC_MAX_STEPS_REACHED = -1
CBC_STEP_SIZE_LIMIT = -2
CBC_CONVERGENCE_FAILURE = -3
CBC_SINGULARITY_OR_NAN = -4

CBC_FIELD_CONST = 0
CBC_FIELD_SW_N = 1
CBC_FIELD_SW_A = 2

CBC_MAP_AUTO = 0
CBC_MAP_LANE = 1
CBC_MAP_GROUP = 2

CBC_LIN_DIRECT = 0

CBC_EVENT_BRACKET = 0
CBC_EVENT_EXACT = 1
EVENT_MODES = {"bracket": CBC_EVENT_BRACKET, "exact": CBC_EVENT_EXACT}

CBC_SERVE_RESET = 1
CBC_SERVE_USE_CONTROL = 2

MATH_CODES = {"sincos": 0, "exp": 1, "log": 2, "pow": 3, "rcp": 4, "sincos_kern": 5, "sincos_tiny": 6,
              "rotate": 7, "pow1p": 8}

FIELD_CODES = {"CONST": CBC_FIELD_CONST, "SW_N": CBC_FIELD_SW_N, "SW_A": CBC_FIELD_SW_A}

_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_f64 = ctypes.c_double
_vp = ctypes.c_void_p


class ModelSpecC(ctypes.Structure):
    _fields_ = [("sd", _i32), ("nup", _i32), ("nsp", _i32), ("ncm", _i32), ("nk", _i32),
                ("ne", _i32), ("ncf", _i32), ("nct", _i32), ("field", _i32)]


class ExecSpecC(ctypes.Structure):
    _fields_ = [("ns", _i32), ("spb", _i32), ("ups", _i32), ("ndo", _i32), ("tile", _i32),
                ("mapping", _i32), ("block_size", _i32)]


class SolverSpecC(ctypes.Structure):
    _fields_ = [("atol", _f64), ("rtol", _f64), ("min_step", _f64), ("max_step", _f64), ("etol", _f64),
                ("lin_atol", _f64), ("lin_rtol", _f64), ("lin_relaxation", _f64),
                ("lin_max_iter", _i32), ("linsolve", _i32), ("col_threshold", _f64), ("event_mode", _i32)]


class BuffersC(ctypes.Structure):
    _fields_ = [(name, _vp) for name in (
        "actual_time", "time_end", "time_begin", "time_step", "actual_state",
        "unit_params", "system_params", "control_params", "coupling_matrices",
        "dense_index", "dense_time", "dense_state", "actual_event", "status_flags",
        "total_steps", "rejected_steps")]


class PeriodSpecC(ctypes.Structure):
    _fields_ = [("period", _f64), ("first_window", _i32), ("n_windows", _i32),
                ("max_r", _vp), ("mean_x", _vp), ("resume", _i32)]


class PlanInfoC(ctypes.Structure):
    _fields_ = [("mapping", _i32), ("block_size", _i32), ("grid_size", _i32), ("systems_per_block", _i32),
                ("lds_bytes", _i32), ("num_regs", _i32), ("local_bytes", _i32), ("max_threads", _i32)]


class ServeSpecC(ctypes.Structure):
    _fields_ = [("capacity", _i32), ("waves", _i32), ("refill", _i32), ("poll_interval", _i32),
                ("max_attempts", _i64), ("idle_us", _i32), ("copy_state", _i32)]


class ServeViewC(ctypes.Structure):
    _fields_ = [("done_seq", _vp), ("out_state", _vp), ("out_misc", _vp), ("stage_state", _vp), ("stage_up", _vp),
                ("stage_cm", _vp), ("capacity", _i32), ("waves", _i32)]


CBC_ENV_MAX_ACTIONS = 16
CBC_ENV_MAX_PARTS = 8
CBC_ENV_MAX_BUBBLES = 16
CBC_ENV_STATIC_MAX_ENVS = 1 << 20   # envs served by the static workspace of cbc_env_post_static
CBC_ENV_PART_STATE = 0
CBC_ENV_PART_TARGET = 1
CBC_ENV_PART_FIXED = 2
CBC_SCENE_RANDOM = 0
CBC_SCENE_EQUIDISTANT = 1
CBC_SCENE_FIXED = 2


class EnvSpecC(ctypes.Structure):
    _fields_ = [("num_envs", _i32), ("num_bubbles", _i32), ("tile", _i32), ("num_control_rows", _i32),
                ("num_actions", _i32), ("action_row", _i32 * CBC_ENV_MAX_ACTIONS),
                ("action_scale", _f64 * CBC_ENV_MAX_ACTIONS), ("num_pa", _i32), ("pa_col", _i32 * CBC_ENV_MAX_ACTIONS),
                ("num_parts", _i32), ("part_kind", _i32 * CBC_ENV_MAX_PARTS), ("part_row", _i32 * CBC_ENV_MAX_PARTS),
                ("part_stack", _i32 * CBC_ENV_MAX_PARTS), ("part_hist", _i32 * CBC_ENV_MAX_PARTS),
                ("obs_dim", _i32), ("history_floats", _i64), ("x_part", _i32), ("target_part", _i32),
                ("apply_termination", _i32), ("episode_length", _i64), ("position_tolerance", ctypes.c_float),
                ("r_dnorm", ctypes.c_float), ("reward_exp", ctypes.c_float), ("w_distance", ctypes.c_float),
                ("w_intensity", ctypes.c_float), ("positive_terminal_reward", ctypes.c_float),
                ("negative_terminal_reward", ctypes.c_float), ("next_time_step", _f64),
                ("env_begin", _i32), ("env_count", _i32)]


class EnvBuffersC(ctypes.Structure):
    _fields_ = [(name, _vp) for name in (
        "actual_state", "actual_time", "time_step", "actual_event", "status_flags", "control_params",
        "actions", "history", "observations", "obs_bias", "obs_rscale", "rewards", "total_rewards",
        "algo_steps", "failure", "success", "terminated", "time_outs", "done_ids", "workspace")]


class SceneSideC(ctypes.Structure):
    _fields_ = [("method", _i32), ("random_align", _i32), ("base", _f64 * CBC_ENV_MAX_BUBBLES), ("offset", _f64),
                ("max_start", _f64), ("u_gaps", _vp), ("u_align", _vp)]


class EnvSceneC(ctypes.Structure):
    _fields_ = [("lo", _f64), ("available", _f64), ("min_distance", _f64), ("room_per_gap", _f64),
                ("extra_gap_limit", _f64), ("bubbles", SceneSideC), ("targets", SceneSideC)]


class EnvCountsC(ctypes.Structure):
    _fields_ = [("done", _i32), ("failures", _i32), ("successes", _i32), ("time_outs", _i32)]


class TrajSpecC(ctypes.Structure):
    _fields_ = [("num_envs", _i32), ("num_bubbles", _i32), ("tile", _i32), ("ndo", _i32), ("num_saved", _i32),
                ("saved_rows", _i32 * 4), ("obs_dim", _i32), ("num_actions", _i32), ("max_steps", _i32),
                ("max_points", _i64), ("step_length", _f64)]


class TrajBuffersC(ctypes.Structure):
    _fields_ = [(name, _vp) for name in (
        "observations", "actions", "rewards", "dense_index", "dense_time", "dense_state",
        "obs", "act", "rew", "time", "states", "steps", "points", "overflow")]


class TrajPackedC(ctypes.Structure):
    _fields_ = [("obs", _vp), ("act", _vp), ("rew", _vp), ("time", _vp), ("states", _vp), ("total_points", _i64)]


# name -> (restype, argtypes); every entry point of include/cbc.h
SIGNATURES = {
    "cbc_abi_version": (_i32, []),
    "cbc_last_error": (ctypes.c_char_p, []),
    "cbc_plan_create": (_i32, [ctypes.POINTER(ModelSpecC), ctypes.POINTER(ExecSpecC),
                               ctypes.POINTER(SolverSpecC), ctypes.POINTER(_vp)]),
    "cbc_plan_query": (_i32, [_vp, ctypes.POINTER(PlanInfoC)]),
    "cbc_plan_destroy": (None, [_vp]),
    "cbc_launch": (_i32, [_vp, ctypes.POINTER(BuffersC), _i32, _vp]),
    "cbc_solve": (_i32, [_vp, ctypes.POINTER(BuffersC), _i32, _i64, _vp,
                         ctypes.POINTER(_f64), _i32, ctypes.POINTER(_i32)]),
    "cbc_solve_periodic": (_i32, [_vp, ctypes.POINTER(BuffersC), ctypes.POINTER(PeriodSpecC), _i64, _vp]),
    "cbc_solve_queued": (_i32, [_vp, ctypes.POINTER(BuffersC), _vp, _i32, _i32, _i32, _i32, _i64, _vp,
                                ctypes.POINTER(_i32)]),
    "cbc_launch_queued": (_i32, [_vp, ctypes.POINTER(BuffersC), _vp, _i32, _i32, _i32, _i32, _vp]),
    "cbc_launch_fused": (_i32, [ctypes.POINTER(_vp), ctypes.POINTER(ctypes.POINTER(BuffersC)), ctypes.POINTER(_vp),
                                ctypes.POINTER(_i32), ctypes.POINTER(_i32), _i32, _i32, _i32, _vp]),
    "cbc_fused_supported": (_i32, [ctypes.POINTER(_i32), _i32, _i32, _i32]),
    "cbc_plan_active_count": (_i32, [_vp, _vp, ctypes.POINTER(_i32)]),
    "cbc_serve_create": (_i32, [_vp, ctypes.POINTER(BuffersC), ctypes.POINTER(ServeSpecC), _vp, ctypes.POINTER(_vp)]),
    "cbc_serve_get_view": (_i32, [_vp, ctypes.POINTER(ServeViewC)]),
    "cbc_serve_submit": (_i32, [_vp, _vp, _vp, _i32, _f64, _vp, _vp, _vp]),
    "cbc_serve_poll": (_i32, [_vp, ctypes.POINTER(_i32)]),
    "cbc_serve_stats": (_i32, [_vp, ctypes.POINTER(_i64), ctypes.POINTER(ctypes.c_uint32),
                               ctypes.POINTER(ctypes.c_uint32)]),
    "cbc_serve_destroy": (_i32, [_vp]),
    "cbc_selftest_math": (_i32, [_i32, _vp, _vp, _vp, _i32, _vp]),
    "cbc_selftest_solve": (_i32, [_i32, _vp, _vp, _i32, _vp]),
    "cbc_env_workspace_words": (_i64, [_i32]),
    "cbc_env_set_action": (_i32, [ctypes.POINTER(EnvSpecC), ctypes.POINTER(EnvBuffersC), _vp]),
    "cbc_env_post": (_i32, [ctypes.POINTER(EnvSpecC), ctypes.POINTER(EnvBuffersC), _i64, _vp,
                            ctypes.POINTER(EnvCountsC)]),
    "cbc_env_post_static": (_i32, [ctypes.POINTER(EnvSpecC), ctypes.POINTER(EnvBuffersC), _i64, _vp,
                                   ctypes.POINTER(EnvCountsC)]),
    "cbc_env_reset": (_i32, [ctypes.POINTER(EnvSpecC), ctypes.POINTER(EnvBuffersC), ctypes.POINTER(EnvSceneC),
                             _vp, _i32, _i64, _vp, _vp, _vp, _vp]),
    "cbc_env_scene_positions": (_i32, [ctypes.POINTER(EnvSceneC), _i32, _i32, _vp, _vp, _vp]),
    "cbc_traj_append": (_i32, [ctypes.POINTER(TrajSpecC), ctypes.POINTER(TrajBuffersC), _vp]),
    "cbc_traj_gather": (_i32, [ctypes.POINTER(TrajSpecC), ctypes.POINTER(TrajBuffersC), _vp, _i32, _vp, _vp,
                               ctypes.POINTER(TrajPackedC), _vp]),
}


class CBCError(RuntimeError):
    """A libcbc call returned a negative status; carries the code and message."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"libcbc error {code}: {msg}")
        self.code = code


_lib = None


def lib() -> ctypes.CDLL:
    """The loaded library (loaded and checked on first use)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"libcbc.so not found at {LIB_PATH}; build it with "
            "`python coupled-bubble-control_amd/build_lib.py` (or __graft_entry__.build())")
    handle = ctypes.CDLL(str(LIB_PATH), mode=ctypes.RTLD_GLOBAL)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(handle, name)
        fn.restype = res
        fn.argtypes = args
    v = handle.cbc_abi_version()
    if v != CBC_ABI_VERSION:
        raise RuntimeError(f"libcbc ABI version {v} != expected {CBC_ABI_VERSION}")
    _lib = handle
    return _lib


def check(rc: int) -> int:
    """Raise CBCError for a negative return code."""
    if rc < 0:
        msg = lib().cbc_last_error().decode(errors="replace")
        raise CBCError(rc, msg)
    return rc


def ptr(t) -> int | None:
    """Device pointer of a torch tensor (None for None)."""
    if t is None:
        return None
    return t.data_ptr()


def stream_handle(stream=None) -> int | None:
    """hipStream_t of a torch stream (current stream when None)."""
    if stream is None:
        stream = torch.cuda.current_stream()
    return stream.cuda_stream or None
