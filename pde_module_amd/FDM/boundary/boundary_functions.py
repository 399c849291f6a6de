"""Boundary values (reference ``pm/FDM/boundary/boundary_functions.py``).

The reference passes Warp functions evaluated per boundary node.  The two it
ships are a constant (``get_constant_func``, :9-47) and a linear ramp in time
(``get_ramp_func``, :51-92); both are represented here as plain records the
boundary kernel evaluates on the GPU: the constant is rounded to FP32, the ramp
is ``min(t / T_max, 1) * constant`` in FP32 with ``t`` read from device memory at
launch (so a captured graph sees the current time)."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class ConstantValue:
    constant: float

    @property
    def value32(self):
        return np.float32(self.constant)

    tmax32 = np.float32(0.0)


@dataclass(frozen=True)
class RampValue:
    constant: float
    T_max: float

    def __post_init__(self):
        if not np.float32(self.T_max) > 0:
            raise ValueError("T_max must be positive")

    @property
    def value32(self):
        return np.float32(self.constant)

    @property
    def tmax32(self):
        return np.float32(self.T_max)


def get_constant_func(constant: float, *args, **kwargs) -> ConstantValue:
    return ConstantValue(float(constant))


def get_ramp_func(constant: float, T_max: float, *args, **kwargs) -> RampValue:
    return RampValue(float(constant), float(T_max))


def as_boundary_value(value):
    """int/float -> ConstantValue; Constant/RampValue pass through; anything
    else (arbitrary per-node functions) is not supported by the GPU tables."""
    if isinstance(value, (ConstantValue, RampValue)):
        return value
    if isinstance(value, (bool, np.bool_)):
        raise TypeError("boundary value must be a number")
    if isinstance(value, (int, float, np.integer, np.floating)):
        return ConstantValue(float(value))
    raise TypeError(
        "boundary values are constants, get_constant_func(...) or get_ramp_func(...); "
        f"arbitrary functions are not supported (got {type(value).__name__})")
