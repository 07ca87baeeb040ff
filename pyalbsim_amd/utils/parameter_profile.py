"""Depth-profile parameters accepted by the water-layer configuration.

Mirrors the ScalarParameter protocol of the reference
(src/alb_sim/utils/parameter_profile.py:9-99): a layer property is either a
plain float or an object with ``at(t)`` evaluated at the fractional depth
``t`` in [0, 1] of the layer.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Protocol, Union

import numpy as np


class ScalarParameter(Protocol):
    def at(self, t): ...


NumberOrScalar = Union[float, ScalarParameter]


def normalize_number_or_scalar(value):
    """Ints become floats, profile objects pass through, bools are rejected.

    Same contract as parameter_profile.py:22-43 (TypeError for bool).
    """
    if isinstance(value, bool):
        raise TypeError("Boolean is not a valid numeric parameter")
    if isinstance(value, int):
        return float(value)
    return value


@dataclass(frozen=True)
class LinearParameter:
    """start + t (end - start), t as float32 (parameter_profile.py:47-68)."""

    start: float
    end: float

    def at(self, t):
        scalar = np.isscalar(t)
        tt = np.asarray(t, dtype=np.float32)
        out = self.start + tt * (self.end - self.start)
        return out.item() if scalar else out


@dataclass(frozen=True)
class ExponentialParameter:
    """exp(log start + t (log end - log start)) (parameter_profile.py:71-99)."""

    start: float
    end: float

    def __post_init__(self):
        if self.start <= 0 or self.end <= 0:
            raise ValueError("ExponentialProfile requires positive values")

    def at(self, t):
        a = np.log(self.start)
        b = np.log(self.end)
        return np.exp(a + t * (b - a))
