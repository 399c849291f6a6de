"""Grad module (reference ``pm/FDM/module/grad.py:10-88``).

The reference's kernel factory calls an undefined helper at this snapshot
(``pm/FDM/kernel/grad.py:33``), so every reference Grad raises NameError at
setup; this module implements the semantics of its ``create_grad_op``
(``grad.py:56-102``), which is complete."""
from __future__ import annotations

import numpy as np

from ...stencil import setup
from ..functional import grad
from ..kernel import create_Grad_kernel
from .explicitUniformGridStencil import ExplicitUniformGridStencil


def central_weights(dx):
    """(-1/(2dx), 0, 1/(2dx)) in FP64, rounded to FP32 (reference
    ``module/grad.py:37-39``)."""
    return tuple(np.float32(v) for v in (-1.0 / (2 * dx), 0.0, 1.0 / (2 * dx)))


class Grad(ExplicitUniformGridStencil):
    """Central gradient.  vec1 input -> vec(D) output (or a (1, D) matrix with
    ``force_matrix``); vec(N) input -> (N, D) matrix output."""

    def __init__(self, inputs: int, grid_shape, dx: float, ghost_cells: int, stencil=None,
                 force_matrix: bool = False, float_dtype=np.float32) -> None:
        self.force_matrix = force_matrix
        dimension = self.calculate_dimension_from_grid_shape(grid_shape)
        if stencil is not None:
            raise ValueError("Custom stencil not implemented yet")
        self.stencil = central_weights(dx)
        if not (type(inputs) is int and inputs > 0):
            raise AssertionError("inputs must be a positive int")
        output_shape = dimension if (inputs == 1 and not force_matrix) else (inputs, dimension)
        if dimension > 3:
            raise AssertionError("dimension <= 3")
        super().__init__(inputs, output_shape, dx, ghost_cells, float_dtype)

    @setup
    def initialize_kernel(self, input_array, *args, **kwargs) -> None:
        self._check_input(input_array)
        if len(self.inputs) != 1:
            raise AssertionError("Grad Only For Vectors")
        self.kernel = create_Grad_kernel(self.input_dtype, self.output_dtype,
                                         tuple(input_array.shape[:3]), self.stencil,
                                         self.ghost_cells)
        self.kernel_dim = self.grid_shape_with_no_ghost_cells(tuple(input_array.shape[:3]),
                                                              self.ghost_cells)

    def forward(self, input_array, alpha: float = 1.0):
        return grad(self.kernel, input_array, alpha, self.output_array, self.ghost_cells,
                    device=self.device)
