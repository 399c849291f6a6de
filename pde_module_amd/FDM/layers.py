"""Host side of the absorbing layers (sponge and viscous damping): layer node
sets, distance levels and the damping profile (reference
``pm/FDM/module/farfield.py:58-89``, ``viscousDampingLayer.py:56-87`` for the
node sets and ``pm/FDM/kernel/farfield.py:36-77``,
``viscousDampingLayer.py:36-74`` for the per-node factor).

The reference computes, on the device and for every layer node, the axis of
the nearest face (``argmin`` over the eligible axes of min(i, s - 1 - i),
lowest axis on ties), the distance level ``dx`` from the layer's inner limit
and ``sigma_max * (float(dx) / float(num_layers)) ** p`` in FP32.  Here the
level of each node is resolved once on the host and the profile
``(dx / num_layers) ** p`` per level is a small table (FP32 quotient, power in
FP64, rounded once), so the device multiplies ``scale * profile[level]``.
The reference's FP32 ``**`` (the device power function) agrees with the
table to its own error bound of a couple of ulp.
"""
from __future__ import annotations

import numpy as np
import torch

from .. import _native as N

INT32_MAX = np.iinfo(np.int32).max


def layer_groups(grid_shape, num_layers):
    """C-order node indices of the padded grid and the boolean group masks
    ``-X, +X, -Y, ...`` (nodes within num_layers of that face) of its eligible
    axes (reference ``farfield.py:64-74``)."""
    indices = np.moveaxis(np.indices(tuple(grid_shape), dtype=np.int32), 0, -1).reshape(-1, 3)
    groups = {}
    for i, (s, name) in enumerate(zip(grid_shape, "XYZ")):
        if s > 1:
            groups[f"-{name}"] = indices[:, i] < num_layers
            groups[f"+{name}"] = indices[:, i] >= (s - num_layers)
    return indices, groups


def layer_points(indices, groups, excluded):
    """Nodes of every non-excluded group, in C order (``farfield.py:76-89``)."""
    masks = [m for k, m in groups.items() if k not in excluded]
    if not masks:
        raise AssertionError("every layer group is excluded")
    mask = masks[0]
    for m in masks[1:]:
        mask = np.logical_or(mask, m)
    return indices[mask]


def layer_levels(points, grid_shape, num_layers):
    """Distance level of each layer node: num_layers minus its distance to
    the nearest face (the reference kernel's ``dx``, ``farfield.py:66-72``)."""
    gs = np.asarray(grid_shape, dtype=np.int64)
    pts = np.asarray(points, dtype=np.int64)
    dist = np.where(gs > 1, np.minimum(pts, gs - pts - 1), INT32_MAX)
    axis = np.argmin(dist, axis=1)
    levels = num_layers - dist[np.arange(len(pts)), axis]
    if len(levels) and (levels.min() < 1 or levels.max() > num_layers):
        raise AssertionError("layer node outside the layer")
    return levels.astype(np.int32)


def layer_profile(num_layers, p):
    """prof[d - 1] = fp32((fp32(d) / fp32(num_layers)) ** fp32(p)), d = 1..num_layers."""
    d = np.arange(1, num_layers + 1, dtype=np.float32)
    base = d / np.float32(num_layers)
    return np.power(base.astype(np.float64), np.float64(np.float32(p))).astype(np.float32)


class LayerKernel:
    """The counterpart of the reference's ``create_spongeLayer_kernel`` /
    ``create_dampingLayer_kernel`` products: element shape, grid, layer
    thickness and the profile table (uploaded once per device)."""

    def __init__(self, op, input_dtype, num_layers, p, grid_shape, ghost_cells):
        self.op = op
        self.input_dtype = tuple(input_dtype)
        self.ncomp = int(np.prod(self.input_dtype))
        if self.ncomp > N.LAYER_MAXC:
            raise NotImplementedError(f"elements of more than {N.LAYER_MAXC} floats")
        self.num_layers = int(num_layers)
        if self.num_layers < 1:
            raise ValueError("num_layers must be >= 1")
        self.p = np.float32(p)
        self.grid_shape = tuple(int(s) for s in grid_shape)
        self.ghost_cells = ghost_cells
        self.profile = layer_profile(self.num_layers, self.p)
        self._dev = {}

    def profile_on(self, device):
        t = self._dev.get(device)
        if t is None:
            t = torch.from_numpy(self.profile).to(device)
            self._dev[device] = t
        return t


def point_table(points, grid_shape, num_layers, device):
    """Device table of ``pde_layer_point`` records (node, level) for the
    layer nodes ``points`` ((M, 3) grid indices)."""
    pts = np.asarray(points, dtype=np.int64).reshape(-1, 3)
    gs = tuple(int(s) for s in grid_shape)
    rec = np.zeros(len(pts), N.LAYER_POINT_DTYPE)
    rec["node"] = (pts[:, 0] * gs[1] + pts[:, 1]) * gs[2] + pts[:, 2]
    rec["level"] = layer_levels(pts, gs, num_layers)
    return torch.from_numpy(rec.view(np.uint8)).to(device)


def launch_layer(kernel, x, table, farfield_condition, scale, out):
    """One ``pde_damping_layer_f32`` launch over the table's nodes."""
    N.check_field(x, "input_array")
    N.check_field(out, "output_array")
    if tuple(x.shape[3:]) != kernel.input_dtype or out.shape != x.shape:
        raise AssertionError("input / output must match the kernel's element shape")
    if tuple(x.shape[:3]) != kernel.grid_shape:
        raise AssertionError("field grid differs from the layer's grid_shape")
    npts = table.numel() // N.LAYER_POINT_DTYPE.itemsize
    ff = None
    if farfield_condition is not None:
        ffv = np.asarray(farfield_condition.detach().cpu() if isinstance(farfield_condition, torch.Tensor)
                         else farfield_condition, dtype=np.float32).reshape(-1)
        if ffv.size == 1 and kernel.ncomp > 1:
            ffv = np.repeat(ffv, kernel.ncomp)
        if ffv.size != kernel.ncomp:
            raise AssertionError(f"farfield_condition needs {kernel.ncomp} values, got {ffv.size}")
        ff = np.ascontiguousarray(ffv)
    prof = kernel.profile_on(x.device)
    N.call("pde_damping_layer_f32", N.dev_ptr(x), N.dev_ptr(out), table.data_ptr(), npts,
           prof.data_ptr(), kernel.num_layers, kernel.ncomp,
           ff.ctypes.data if ff is not None else None, np.float32(scale), N.stream_of(x))
    return out
