"""Kernel descriptors (the counterpart of the reference's kernel factories in
``pm/FDM/kernel/{laplacian,grad,divergence}.py``).

The reference JIT-builds one Warp kernel per (dtype, grid shape, stencil,
ghost count).  Here the HIP kernels are compiled once for every layout, so a
"kernel" is the small host record of what one launch needs: the FP32 stencil
weights, the element shapes and the ghost width, validated up front with the
reference's assertions.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .utils import eligible_dims_and_shift


@dataclass(frozen=True)
class StencilKernel:
    op: str                 # "laplacian" | "grad" | "divergence"
    weights: tuple          # three float32 weights (s = -1, 0, +1)
    in_shape: tuple         # element shape of the input, (C,) or (N, D)
    out_shape: tuple        # element shape of the output
    grid_shape: tuple       # padded (X, Y, Z)
    ghost_cells: int

    @property
    def dimension(self) -> int:
        return len(eligible_dims_and_shift(self.grid_shape, self.ghost_cells)[0])


def _check_stencil(stencil):
    w = tuple(np.float32(x) for x in stencil)
    if len(w) % 2 != 1:
        raise AssertionError("the stencil needs an odd number of weights")
    if len(w) != 3:
        raise ValueError("only the default central stencil is supported")
    return w


def create_Laplacian_kernel(input_shape, grid_shape, stencil, ghost_cells):
    """Reference ``pm/FDM/kernel/laplacian.py:7-61``: vector fields only."""
    if len(input_shape) != 1:
        raise AssertionError("the input must be a vector field")
    return StencilKernel("laplacian", _check_stencil(stencil), tuple(input_shape),
                         tuple(input_shape), tuple(grid_shape), ghost_cells)


def create_Grad_kernel(input_shape, output_shape, grid_shape, stencil, ghost_cells):
    """Reference ``pm/FDM/kernel/grad.py:8-102`` (and ``create_grad_op:56-102``)."""
    dims, _ = eligible_dims_and_shift(grid_shape, ghost_cells)
    if len(input_shape) != 1:
        raise AssertionError("the gradient takes vector fields")
    if len(output_shape) == 1:
        if input_shape[0] != 1:
            raise AssertionError("a vector-valued gradient needs a scalar (length-1) input; use the matrix form")
        if output_shape[0] != len(dims):
            raise AssertionError("vector output length must equal the grid dimension")
    else:
        N, D = output_shape
        if D != len(dims) or N != input_shape[0]:
            raise AssertionError("matrix output must be (inputs, dimension)")
    return StencilKernel("grad", _check_stencil(stencil), tuple(input_shape),
                         tuple(output_shape), tuple(grid_shape), ghost_cells)


def create_Divergence_kernel(input_shape, output_shape, grid_shape, stencil, ghost_cells):
    """Reference ``pm/FDM/kernel/divergence.py:7-108``."""
    dims, _ = eligible_dims_and_shift(grid_shape, ghost_cells)
    if len(input_shape) == 2:
        if input_shape[1] != len(dims):
            raise AssertionError("the matrix must have one column per grid dimension")
    elif input_shape[0] != len(dims):
        raise AssertionError("vector input length must equal the grid dimension")
    return StencilKernel("divergence", _check_stencil(stencil), tuple(input_shape),
                         tuple(output_shape), tuple(grid_shape), ghost_cells)


def create_spongeLayer_kernel(input_dtype, num_layers, beta, grid_shape, ghost_cells):
    """Reference ``pm/FDM/kernel/farfield.py:8-78``."""
    from .layers import LayerKernel
    return LayerKernel("sponge", input_dtype, num_layers, beta, grid_shape, ghost_cells)


def create_dampingLayer_kernel(input_dtype, num_layers, p, grid_shape, ghost_cells):
    """Reference ``pm/FDM/kernel/viscousDampingLayer.py:8-76``."""
    from .layers import LayerKernel
    return LayerKernel("damping", input_dtype, num_layers, p, grid_shape, ghost_cells)
