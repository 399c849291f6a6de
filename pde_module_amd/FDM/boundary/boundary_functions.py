"""Boundary values (reference ``pm/FDM/boundary/boundary_functions.py`` and the
user boundary functions of ``pm/FDM/boundary/functional_boundary.py:139-175``
and ``pm/deprecated/func_boundary.py:10-136``).

The reference passes Warp functions evaluated per boundary node.  Here a value
is one of:

* ``ConstantValue`` (``get_constant_func``, reference :9-47): the constant
  rounded to FP32;
* ``RampValue`` (``get_ramp_func``, reference :51-92): ``min(t / T_max, 1) *
  constant`` in FP32;
* ``SineValue`` (``get_sine_func``): ``amplitude * sin(omega * t)`` in FP32,
  the examples' ``sin_wave`` (``examples/wave2d/Interference.py:53-62``);
* ``FunctionValue``: any Python callable with the reference's argument list,
  evaluated on the device by the boundary module before each application.

The first three are records the boundary kernels evaluate themselves, with
``t`` read from device memory at launch (so a captured graph sees the current
time).  A ``FunctionValue`` is called once per application with tensors for a
whole group at once instead of once per node:

* functional ``GridBoundary``: ``func(current_values, nodeID, coordinates, t,
  dx, params)`` returns a value broadcastable to ``(M, C)`` (the node vector,
  as the reference's function returns a ``vec``);
* legacy ``GridBoundary``: ``func(current_values, nodeID, varID, coordinates,
  t, dx, params)`` returns a value broadcastable to ``(M,)`` (component
  ``varID``);

where ``current_values`` is the boundary's input field ``(X, Y, Z, C)`` (the
values before the boundary is applied), ``nodeID`` the group's ``(M, 3)``
int64 node indices on the field's device, ``coordinates`` the
``grid_coordinates`` given to the boundary as a device tensor ``(X, Y, Z, 3)``
(or None), ``t`` a 0-d float32 device tensor, ``dx`` the FP32 grid spacing and
``params`` the call's ``params[group_name]`` (0 when absent, as the reference's
default).  The results are written into the device time array the kernels
read (``t[1 + slot]``), so the function's torch operations run on the GPU, in
the same stream, and are captured with the step when it is graph-captured (a
function that synchronises, e.g. a Python ``if`` on ``t``, works eagerly but
cannot be captured).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Callable

import numpy as np

from ..._native import BCV_ARRAY, BCV_CONST, BCV_SINE


@dataclass(frozen=True)
class ConstantValue:
    constant: float
    vkind = BCV_CONST

    @property
    def value32(self):
        return np.float32(self.constant)

    tmax32 = np.float32(0.0)


@dataclass(frozen=True)
class RampValue:
    constant: float
    T_max: float
    vkind = BCV_CONST

    def __post_init__(self):
        if not np.float32(self.T_max) > 0:
            raise ValueError("T_max must be positive")

    @property
    def value32(self):
        return np.float32(self.constant)

    @property
    def tmax32(self):
        return np.float32(self.T_max)


@dataclass(frozen=True)
class SineValue:
    """``amplitude * sin(omega * t)``: FP32 product omega * t, the device sine
    of that, times the amplitude."""
    amplitude: float
    omega: float
    vkind = BCV_SINE

    def __post_init__(self):
        if not (np.isfinite(np.float32(self.amplitude)) and np.isfinite(np.float32(self.omega))):
            raise ValueError("amplitude and omega must be finite")

    @property
    def value32(self):
        return np.float32(self.amplitude)

    @property
    def tmax32(self):
        return np.float32(self.omega)


@dataclass(frozen=True, eq=False)
class FunctionValue:
    """A Python boundary function evaluated on the device (see the module
    docstring for its arguments)."""
    func: Callable
    vkind = BCV_ARRAY
    value32 = np.float32(0.0)
    tmax32 = np.float32(0.0)


def get_constant_func(constant: float, *args, **kwargs) -> ConstantValue:
    return ConstantValue(float(constant))


def get_ramp_func(constant: float, T_max: float, *args, **kwargs) -> RampValue:
    return RampValue(float(constant), float(T_max))


def get_sine_func(amplitude: float, omega: float, *args, **kwargs) -> SineValue:
    return SineValue(float(amplitude), float(omega))


def as_boundary_value(value):
    """int/float -> ConstantValue; value records pass through; any other
    callable -> FunctionValue."""
    if isinstance(value, (ConstantValue, RampValue, SineValue, FunctionValue)):
        return value
    if isinstance(value, (bool, np.bool_)):
        raise TypeError("boundary value must be a number or a function")
    if isinstance(value, (int, float, np.integer, np.floating)):
        return ConstantValue(float(value))
    if callable(value):
        return FunctionValue(value)
    raise TypeError("boundary values are numbers, get_constant_func / get_ramp_func / "
                    f"get_sine_func records or callables (got {type(value).__name__})")
