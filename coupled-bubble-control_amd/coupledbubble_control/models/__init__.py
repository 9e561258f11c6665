"""CKM1D model, specs and buffers (reference: coupledbubble_control/models/__init__.py)."""
from .buffers import (ActualStateBuffer, BufferOps, BufferSet, DenseOutputBuffer, ParameterBuffer,
                      StatusBuffer)
from .ckm1d import CKM1D
from .model import MaterialProperties, Model, ModelRuntime, water_25c
from .specs import ExecutionSpec, LayoutEntry, ModelSpec, SolverSpec

__all__ = ["CKM1D", "MaterialProperties", "Model", "ModelRuntime", "water_25c", "ExecutionSpec", "LayoutEntry",
           "ModelSpec", "SolverSpec", "ActualStateBuffer", "BufferOps", "BufferSet", "DenseOutputBuffer",
           "ParameterBuffer", "StatusBuffer"]
