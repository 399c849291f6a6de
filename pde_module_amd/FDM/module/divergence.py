"""Divergence module (reference ``pm/FDM/module/divergence.py:11-111``).

The reference launches its kernel twice per call with the same result
(``divergence.py:98-111``); one launch here."""
from __future__ import annotations

from collections.abc import Iterable

import numpy as np

from ...stencil import setup
from ..functional import divergence
from ..kernel import create_Divergence_kernel
from .explicitUniformGridStencil import ExplicitUniformGridStencil
from .grad import central_weights


class Divergence(ExplicitUniformGridStencil):
    """Central divergence: vec(D) input -> vec1; (N, D) matrix input -> vec(N)."""

    def __init__(self, inputs, grid_shape, dx: float, ghost_cells: int, stencil=None,
                 float_dtype=np.float32) -> None:
        if stencil is not None:
            raise ValueError("Custom stencil not implemented yet")
        self.stencil = central_weights(dx)
        dimension = self.calculate_dimension_from_grid_shape(grid_shape)
        is_vector = False
        if isinstance(inputs, Iterable):
            inputs = tuple(inputs)
            if not all(type(i) is int for i in inputs):
                raise AssertionError("inputs must be ints")
            if len(inputs) == 1:
                inputs, is_vector = inputs[0], True
            elif len(inputs) != 2:
                raise AssertionError("matrix inputs are (N, D)")
        elif type(inputs) is int:
            is_vector = True
        else:
            raise ValueError("Inputs must be int or tuple|list of int")
        if is_vector:
            if inputs != dimension:
                raise AssertionError("vector input length must equal the grid dimension")
            outputs = 1
        else:
            if inputs[1] != dimension:
                raise AssertionError("matrix columns must equal the grid dimension")
            outputs = inputs[0]
        super().__init__(inputs, outputs, dx, ghost_cells, float_dtype=float_dtype)

    @setup
    def initialize_kernel(self, input_array, *args, **kwargs) -> None:
        self._check_input(input_array)
        self.kernel = create_Divergence_kernel(self.input_dtype, self.output_dtype,
                                               tuple(input_array.shape[:3]), self.stencil,
                                               self.ghost_cells)
        self.kernel_dim = self.grid_shape_with_no_ghost_cells(tuple(input_array.shape[:3]),
                                                              self.ghost_cells)

    def forward(self, input_array, alpha: float = 1.0, *args, **kwargs):
        return divergence(self.kernel, input_array, alpha, self.ghost_cells, self.output_array,
                          device=self.device)
