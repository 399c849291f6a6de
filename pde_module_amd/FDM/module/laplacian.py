"""Laplacian module (reference ``pm/FDM/module/laplacian.py:10-73``)."""
from __future__ import annotations

import numpy as np

from ...stencil import setup
from ..functional import laplacian, laplacian_periodic
from ..kernel import create_Laplacian_kernel
from .explicitUniformGridStencil import ExplicitUniformGridStencil


def laplacian_weights(dx):
    """(1/dx², -2/dx², 1/dx²) computed in FP64 from the Python dx, then rounded
    to FP32 (reference ``module/laplacian.py:32-35``)."""
    return tuple(np.float32(v) for v in (1.0 / dx**2, -2.0 / dx**2, 1.0 / dx**2))


class Laplacian(ExplicitUniformGridStencil):
    """Second-order central Laplacian of a vec(inputs) field.

    ``Laplacian(inputs, dx, ghost_cells)``; ``lap(u, alpha)`` returns the module's
    fixed output with ``alpha * laplacian(u)`` on the interior nodes (output
    ghosts untouched)."""

    def __init__(self, inputs: int, dx: float, ghost_cells: int, stencil=None,
                 float_dtype=np.float32) -> None:
        if stencil is not None:
            raise ValueError("Custom stencil not implemented yet")
        self.stencil = laplacian_weights(dx)
        super().__init__(inputs, inputs, dx, ghost_cells, float_dtype=float_dtype)

    @setup
    def initialize_kernel(self, input_array, *args, **kwargs) -> None:
        self._check_input(input_array)
        if len(self.inputs) != 1:
            raise AssertionError("Laplacian Only For Vectors")
        self.kernel = create_Laplacian_kernel(self.input_dtype, tuple(input_array.shape[:3]),
                                              self.stencil, self.ghost_cells)
        self.kernel_dim = self.grid_shape_with_no_ghost_cells(tuple(input_array.shape[:3]),
                                                              self.ghost_cells)

    def forward(self, input_array, alpha: float = 1.0, *args, **kwargs):
        return laplacian(self.kernel, input_array, alpha, self.output_array, self.ghost_cells,
                         device=self.device)


class PeriodicLaplacian(Laplacian):
    """``Laplacian`` on a periodic grid (no reference counterpart: the reference
    has no periodic boundary).  ``PeriodicLaplacian(inputs, dx, ghost_cells)``;
    ``lap(u, alpha)`` gives the values ``Laplacian`` gives after
    ``PeriodicBoundary(ghost_cells)(u)``, reading interior values only: u's
    ghost layers are neither read nor written."""

    def forward(self, input_array, alpha: float = 1.0, *args, **kwargs):
        return laplacian_periodic(self.kernel, input_array, alpha, self.output_array,
                                  self.ghost_cells, device=self.device)
