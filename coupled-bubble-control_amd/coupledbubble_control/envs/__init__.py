"""Environment-side adapters of the solver path (reference: coupledbubble_control/envs)."""
from .served import ServedVecEnv

__all__ = ["ServedVecEnv"]
