"""Functional tier: one launch over the interior of a padded field (reference
``pm/FDM/functional/_laplacian.py:5-36``, ``_grad.py:5-36``,
``_divergence.py:4-24``; same argument order, including divergence's
``ghost_cells`` before ``output_array``).

Each call launches asynchronously on torch's current stream of the input's
device and returns ``output_array``.  ``device`` is accepted for signature
compatibility; the launch device is the tensors' device.
"""
from __future__ import annotations

import numpy as np
import torch

from .. import _native as N
from .kernel import StencilKernel


def _grid(x: torch.Tensor):
    if x.dim() < 4:
        raise ValueError("fields are (X, Y, Z, C) or (X, Y, Z, N, D) tensors")
    return tuple(int(s) for s in x.shape[:3])


def _check(kernel: StencilKernel, x: torch.Tensor, out: torch.Tensor, op: str):
    if kernel.op != op:
        raise TypeError(f"expected a {op} kernel, got {kernel.op}")
    N.check_field(x, "input_array")
    N.check_field(out, "output_array")
    if tuple(x.shape[3:]) != kernel.in_shape:
        raise AssertionError(f"input element shape {tuple(x.shape[3:])} != {kernel.in_shape}")
    if tuple(out.shape) != tuple(x.shape[:3]) + kernel.out_shape:
        raise AssertionError(f"output shape {tuple(out.shape)} does not match the kernel")
    if out.device != x.device:
        raise ValueError("input and output must be on the same device")


def laplacian(kernel, input_array, alpha, output_array, ghost_cells, device=None):
    _check(kernel, input_array, output_array, "laplacian")
    X, Y, Z = _grid(input_array)
    wm, wc, wp = kernel.weights
    N.call("pde_laplacian_f32", N.dev_ptr(input_array), N.dev_ptr(output_array), X, Y, Z,
           int(kernel.in_shape[0]), int(ghost_cells), wm, wc, wp, np.float32(alpha),
           N.stream_of(input_array))
    return output_array


def laplacian_periodic(kernel, input_array, alpha, output_array, ghost_cells, device=None):
    """``laplacian`` on a periodic grid: every neighbour index that leaves the
    interior wraps to the opposite interior index, so the input's ghost layers
    are neither read nor written (same values as a periodic ghost fill followed
    by ``laplacian``).  No reference counterpart (the reference has no periodic
    boundary)."""
    _check(kernel, input_array, output_array, "laplacian")
    X, Y, Z = _grid(input_array)
    wm, wc, wp = kernel.weights
    N.call("pde_laplacian_periodic_f32", N.dev_ptr(input_array), N.dev_ptr(output_array),
           X, Y, Z, int(kernel.in_shape[0]), int(ghost_cells), wm, wc, wp, np.float32(alpha),
           N.stream_of(input_array))
    return output_array


def grad(kernel, input_array, alpha, output_array, ghost_cells, device=None):
    _check(kernel, input_array, output_array, "grad")
    X, Y, Z = _grid(input_array)
    wm, wc, wp = kernel.weights
    N.call("pde_grad_f32", N.dev_ptr(input_array), N.dev_ptr(output_array), X, Y, Z,
           int(kernel.in_shape[0]), int(ghost_cells), wm, wc, wp, np.float32(alpha),
           N.stream_of(input_array))
    return output_array


def divergence(kernel, input_array, alpha, ghost_cells, output_array, device=None):
    _check(kernel, input_array, output_array, "divergence")
    X, Y, Z = _grid(input_array)
    wm, wc, wp = kernel.weights
    rows = kernel.in_shape[0] if len(kernel.in_shape) == 2 else 1
    N.call("pde_divergence_f32", N.dev_ptr(input_array), N.dev_ptr(output_array), X, Y, Z,
           int(rows), int(ghost_cells), wm, wc, wp, np.float32(alpha), N.stream_of(input_array))
    return output_array


def fused_lap_grad_div(s, v, lap_out, grad_out, div_out, dx, ghost_cells,
                       alpha_lap=1.0, alpha_grad=1.0, alpha_div=1.0):
    """lap(s), grad(s) and div(v) of a scalar and a vec3 field on one 3D grid,
    reading each input once and writing each output once (one launch, or a
    scalar and a vector launch on grids the Divergence ring takes); each
    output equals the corresponding module's."""
    for name, t in (("s", s), ("v", v), ("lap", lap_out), ("grad", grad_out), ("div", div_out)):
        N.check_field(t, name)
    X, Y, Z = _grid(s)
    g3 = (X, Y, Z)
    if (tuple(s.shape) != g3 + (1,) or tuple(v.shape) != g3 + (3,) or
            tuple(lap_out.shape) != g3 + (1,) or tuple(grad_out.shape) != g3 + (3,) or
            tuple(div_out.shape) != g3 + (1,)):
        raise AssertionError("fused pass needs s, lap, div: (X,Y,Z,1) and v, grad: (X,Y,Z,3)")
    from .module.laplacian import laplacian_weights
    from .module.grad import central_weights
    wl, wc = laplacian_weights(dx), central_weights(dx)
    N.call("pde_fused_lap_grad_div_f32", N.dev_ptr(s), N.dev_ptr(v), N.dev_ptr(lap_out),
           N.dev_ptr(grad_out), N.dev_ptr(div_out), X, Y, Z, int(ghost_cells), *wl, *wc,
           np.float32(alpha_lap), np.float32(alpha_grad), np.float32(alpha_div),
           N.stream_of(s))
    return lap_out, grad_out, div_out


def elementWiseOp(kernel, array_A, array_B, output_array, device=None):
    """Apply the element op ``kernel`` (an ``element_ops`` op or its name) to A
    and B into ``output_array`` over every padded node (reference
    ``pm/FDM/functional/_elementWiseOps.py:5-31``)."""
    from ..stencil.elementWise import resolve_element_op
    op = resolve_element_op(kernel)
    N.check_field(array_A, "array_A")
    N.check_field(array_B, "array_B")
    N.check_field(output_array, "output_array")
    want = op.out_shape(tuple(array_A.shape[3:]), tuple(array_B.shape[3:]))
    if tuple(output_array.shape) != tuple(array_A.shape[:3]) + tuple(want) or \
            tuple(array_B.shape[:3]) != tuple(array_A.shape[:3]):
        raise AssertionError("element-wise shapes do not match")
    op.launch(array_A, array_B, output_array)
    return output_array


def linear_combination(output_array, *terms):
    """``output = c0*x0 + c1*x1 + ...`` (up to four ``(c, x)`` terms) in one pass."""
    from ..stencil.element_ops import linear_combination as _lc
    return _lc(output_array, *terms)


def farfield(kernel, input_array, warp_outer_points, farfield_condition, sigma_max,
             output_array, device=None):
    """Sponge layer (reference ``pm/FDM/functional/_farfield.py:5-41``):
    ``-sigma_max * prof * (input - farfield_condition)`` on the layer nodes of
    ``warp_outer_points`` (a device table from ``FDM.layers.point_table``)."""
    from .layers import launch_layer
    if kernel.op != "sponge":
        raise TypeError(f"expected a sponge layer kernel, got {kernel.op}")
    if farfield_condition is None:
        raise ValueError("farfield_condition is required")
    return launch_layer(kernel, input_array, warp_outer_points, farfield_condition, sigma_max,
                        output_array)


def viscousDampingLayer(kernel, du_dt, warp_outer_points, C_max, output_array, device=None):
    """Viscous damping layer (reference
    ``pm/FDM/functional/_viscousDampingLayer.py:5-38``): ``-C_max * prof * du_dt``
    on the layer nodes."""
    from .layers import launch_layer
    if kernel.op != "damping":
        raise TypeError(f"expected a damping layer kernel, got {kernel.op}")
    return launch_layer(kernel, du_dt, warp_outer_points, None, C_max, output_array)
