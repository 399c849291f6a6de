"""Explicit Euler (reference ``pm/time_step/forwardEuler.py:9-81``)."""
from __future__ import annotations

import numpy as np
import torch

from .. import _native as N
from ..stencil import Stencil, setup


class ForwardEuler(Stencil):
    """``euler(y, k, dt)`` returns the module's fixed output holding
    ``y + dt * k`` over every scalar of the array, ghosts included.  Passing the
    module's own output back as ``y`` or ``k`` updates it in place, the aliasing
    the reference's wave loop relies on."""

    def __init__(self, *args, **kwargs) -> None:
        super().__init__()

    @setup
    def initialize_kernel(self, input_array: torch.Tensor, stencil_values: torch.Tensor,
                          dt: float) -> None:
        self.output_array = torch.empty_like(input_array)
        self.size = input_array.numel()
        if not (input_array.shape == stencil_values.shape == self.output_array.shape):
            raise AssertionError("y and k must have the same shape")

    def forward(self, input_array: torch.Tensor, stencil_values: torch.Tensor, dt: float):
        N.check_field(input_array, "input_array")
        N.check_field(stencil_values, "stencil_values")
        if input_array.numel() != self.size or stencil_values.numel() != self.size:
            raise AssertionError("array size changed after setup")
        N.call("pde_forward_euler_f32", N.dev_ptr(input_array), N.dev_ptr(stencil_values),
               N.dev_ptr(self.output_array), self.size, np.float32(dt),
               N.stream_of(input_array))
        return self.output_array
