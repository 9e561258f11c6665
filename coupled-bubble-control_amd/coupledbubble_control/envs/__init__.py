"""Vector environments over the solver (reference: coupledbubble_control/envs).

``PosNBC1D`` is the reference's position-control environment with its step on
the GPU; ``ServedPosNBC1D`` serves that same step block by block over the serving
kernel (rollouts with the policy overlapped); ``ServedVecEnv`` drives the
solver part of the step alone through the serving queue."""
from .bubble_1D_pos_control import PosNBC1D, SceneGenerator
from .served import ServedPosNBC1D, ServedVecEnv

__all__ = ["PosNBC1D", "SceneGenerator", "ServedPosNBC1D", "ServedVecEnv"]
