"""Base of the uniform-grid operator modules (reference
``pm/FDM/module/ExplicitUniformGridStencil.py:9-179``).

Fields are torch tensors of padded grid shape ``(X, Y, Z)`` followed by the
element shape: ``(C,)`` for a vector (the reference's ``vec(C, float32)``) or
``(N, D)`` for a row-major matrix (``mat(N, D)``).
"""
from __future__ import annotations

import numpy as np
import torch

from ...stencil import Stencil, setup, tuplify
from .. import utils


def dtype_from_shape(shape):
    """Element shape of a vector (``int`` or 1-tuple) or matrix (2-tuple), as in
    reference ``pm/utils/utils.py:35-52``."""
    shape = tuplify(shape)
    if not all(isinstance(x, (int, np.integer)) for x in shape):
        raise AssertionError("contents in input/output must be int only")
    if len(shape) not in (1, 2):
        raise AssertionError(
            "AoS stencils input/output can be either vec (length 1) or matrix (length 2)")
    return tuple(int(x) for x in shape)


class ExplicitUniformGridStencil(Stencil):
    output_array: torch.Tensor | None = None

    def __init__(self, inputs, outputs, dx: float, ghost_cells: int = 0,
                 float_dtype=torch.float32) -> None:
        super().__init__()
        if float_dtype not in (torch.float32, np.float32):
            raise NotImplementedError("the HIP kernels compute in float32 only")
        self.dx = float(np.float32(dx))
        self._inputs = tuplify(inputs)
        self._outputs = tuplify(outputs)
        self.ghost_cells = ghost_cells
        self.float_dtype = torch.float32
        self._input_dtype = dtype_from_shape(self._inputs)
        self._output_dtype = dtype_from_shape(self._outputs)

    # Element shapes play the role of the reference's Warp vector/matrix dtypes.
    @property
    def input_dtype(self):
        return self._input_dtype

    @property
    def output_dtype(self):
        return self._output_dtype

    @property
    def inputs(self):
        return self._inputs

    @property
    def outputs(self):
        return self._outputs

    @property
    def input_dtype_shape(self):
        return self._input_dtype

    @property
    def output_dtype_shape(self):
        return self._output_dtype

    @staticmethod
    def calculate_dimension_from_grid_shape(shape) -> int:
        return utils.dimension_of(shape)

    @staticmethod
    def grid_shape_with_no_ghost_cells(grid_shape, ghost_cells):
        return utils.grid_shape_with_no_ghost_cells(grid_shape, ghost_cells)

    @staticmethod
    def get_ghost_shape(grid_shape, ghost_cells):
        return utils.get_ghost_shape(grid_shape, ghost_cells)

    @setup(order=-1)
    def initialize_array(self, input_array: torch.Tensor, *args, **kwargs) -> None:
        """Allocate the (unzeroed) output once."""
        self.output_array = self.create_output_array(input_array, self.output_dtype)

    @setup(order=1)
    def initialize_kernel(self, input_array: torch.Tensor, *args, **kwargs) -> None:
        ...

    def _check_input(self, input_array: torch.Tensor) -> None:
        if tuple(input_array.shape[3:]) != self.input_dtype:
            raise AssertionError(
                f"input element shape {tuple(input_array.shape[3:])} != {self.input_dtype}")

    def forward(self, input_array, *args, **kwargs):
        raise NotImplementedError()
