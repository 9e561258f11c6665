"""Vector spaces of the environment (reference: coupledbubble_control/envs/common/spaces.py:1-117).

Same classes, constructor arguments and sampling as the reference: a space
owns a torch.Generator on its device, seeded once, so ``sample()`` draws the
reference's stream.  ``Box.scale`` maps a value in [low, high] to [-1, 1]
with float32 bias and reciprocal scale; the environment's observation kernel
applies the same two float32 operations on the device.
"""
from __future__ import annotations

from abc import ABC, abstractmethod
from dataclasses import dataclass
from typing import Optional, Tuple, Union

import torch


@dataclass(frozen=True)
class SpaceData:
    shape: Tuple[int, ...]
    dtype: Optional[torch.dtype] = None
    n: Optional[int] = None
    low: Optional[torch.Tensor] = None
    high: Optional[torch.Tensor] = None


class BaseVectorSpace(ABC):
    def __init__(self, device, num_envs: int = 1, seed: Optional[int] = None):
        self.num_envs = num_envs
        self._seed = seed
        self._device = device
        self._generator = torch.Generator(device=device)
        if seed is not None:
            self._generator.manual_seed(seed)

    @property
    @abstractmethod
    def space_data(self) -> SpaceData:
        ...

    @abstractmethod
    def sample(self) -> torch.Tensor:
        ...

    @property
    def shape(self) -> Tuple[int, ...]:
        return (self.num_envs,) + self.space_data.shape


class Discrete(BaseVectorSpace):
    """Integers in [start, n) (spaces.py:39-57)."""

    def __init__(self, device, n: int, num_envs: int = 1, seed: Optional[int] = None, start: int = 0):
        super().__init__(device, num_envs, seed)
        self._n, self._start = n, start
        self._data = SpaceData(shape=(1,), dtype=torch.long, n=n)

    def sample(self, device: Optional[str] = None) -> torch.Tensor:
        return torch.randint(self._start, self._n, self.shape, generator=self._generator,
                             device=device if device is not None else self._device)

    @property
    def space_data(self) -> SpaceData:
        return self._data


class Box(BaseVectorSpace):
    """A box [low, high]^n (spaces.py:59-117)."""

    def __init__(self, device, low: Union[float, torch.Tensor], high: Union[float, torch.Tensor], n: int = 1,
                 dtype: Optional[torch.dtype] = None, num_envs: int = 1, seed: Optional[int] = None):
        super().__init__(device, num_envs, seed)
        dtype = torch.float32 if dtype is None else dtype

        def bound(v):
            if isinstance(v, torch.Tensor):
                return v.to(dtype=dtype), len(v)
            if isinstance(v, (int, float)):
                return torch.full((n,), v, dtype=dtype), n
            raise TypeError(f"Unsupported type for bound: {type(v)}")

        low, n_low = bound(low)
        high, n_high = bound(high)
        if n_low != n_high:
            raise RuntimeError(f"Bound size mismatch: 'low' has size {n_low}, but 'high' has size {n_high}. "
                               "Both bounds must have the same number of dimensions.")
        self._data = SpaceData(shape=(n_high,), dtype=dtype, n=n, low=low, high=high)
        self._scale = (high - low) * 0.5
        self._rscale = 1.0 / self._scale
        self._bias = (high + low) * 0.5

    def sample(self, device: Optional[str] = None) -> torch.Tensor:
        dev = device if device is not None else self._device
        low, high = self._data.low.to(dev), self._data.high.to(dev)
        u = torch.rand(size=self.shape, dtype=self._data.dtype, generator=self._generator, device=dev)
        return u * (high - low) + low

    def scale(self, x: torch.Tensor) -> torch.Tensor:
        return (x - self._bias.to(x.device)) * self._rscale.to(x.device)

    @property
    def space_data(self) -> SpaceData:
        return self._data
