"""Vector environments over the solver (reference: coupledbubble_control/envs).

``PosNBC1D`` is the reference's position-control environment with its step on
the GPU; ``ServedVecEnv`` drives the same step through the serving queue."""
from .bubble_1D_pos_control import PosNBC1D, SceneGenerator
from .served import ServedVecEnv

__all__ = ["PosNBC1D", "SceneGenerator", "ServedVecEnv"]
