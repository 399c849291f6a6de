"""Staircase immersed boundary (reference
``pm/FDM/boundary/immersedBoundary.py:16-211``).

Solid regions come from boolean functions of the node coordinates
(``from_bool_func``, which adds to an int8 bitmask: nonzero = solid).
``finalize`` finds, on the host, the solid boundary nodes (a solid node with a
fluid neighbour along some eligible axis), the interior solid nodes, each
boundary node's fluid neighbours and the fluid nodes next to the solid, and
creates the (node, component) type/value tables of the legacy ``Boundary``
base (every entry Dirichlet 0 until set; group ``ALL``).  A call copies the
field to the module's output and launches one HIP kernel that writes
Dirichlet values and Neumann running means on the solid boundary nodes and
``fill_value`` on the interior solid nodes.

The reference reads the bitmask (and the field) at index +-1 of every solid
node without a bounds check, so on the outermost layer of the padded grid the
neighbour is the node at flat index +- the axis stride (for a row index -1,
the previous row's last node), as its own slot example relies on for solids
that reach into the ghost rows.  Neighbours are resolved the same way here;
one whose flat index falls outside the array raises ``ValueError`` in
``finalize``.
"""
from __future__ import annotations

from typing import Callable

import numpy as np
import torch

from ... import _native as N
from ...stencil import before_forward, setup
from .boundary import Boundary
from .flags import FLUID_CELL


class ImmersedBoundary(Boundary):
    """``ImmersedBoundary(field, dx, ghost_cells)``; ``from_bool_func(fn, meshgrid)``,
    ``finalize()``, then ``dirichlet_BC('ALL', v)`` / ``vonNeumann_BC('ALL', v)``;
    ``ib(u, fill_value=0.0)`` returns the fixed output."""

    def __init__(self, field: torch.Tensor, dx: float, ghost_cells: int) -> None:
        super().__init__(field, dx, ghost_cells)
        self.bitmask = np.zeros(self.grid_shape, dtype=np.int8)
        self._ops_dev = None

    # ------------------------------------------------------------ geometry
    def from_bool_func(self, fn: Callable, meshgrid) -> None:
        """Add ``fn(*meshgrid)`` (true inside the solid) to the bitmask
        (``immersedBoundary.py:40-52``)."""
        if tuple(np.shape(meshgrid[0])[:3]) != self.grid_shape:
            raise AssertionError("meshgrid must have the field's grid shape")
        mask = np.asarray(fn(*meshgrid))
        self.bitmask += mask.reshape(self.grid_shape).astype(np.int8)

    def show_bitmask(self):
        """The 2D bitmask as an array (plotted when matplotlib is available)."""
        if self.dimension != 2:
            raise AssertionError("Currently only displaying 2D bitmasks supported")
        img = np.array(self.bitmask).squeeze().T
        try:
            from matplotlib import pyplot as plt
        except ImportError:
            return img
        fig, ax = plt.subplots()
        ax.imshow(img, origin="lower", cmap="gray", vmin=0, vmax=1)
        ax.set_title("BitMask")
        plt.show()
        return img

    def _eligible(self):
        return [a for a, s in enumerate(self.grid_shape) if s > 1]

    def _flat(self, idx):
        X, Y, Z = self.grid_shape
        idx = idx.astype(np.int64)
        return (idx[:, 0] * Y + idx[:, 1]) * Z + idx[:, 2]

    def _neighbour_bits(self, idx, axis, shift):
        """Bitmask values at flat index(idx) + shift * stride(axis) (idx: (M, 3))."""
        X, Y, Z = self.grid_shape
        stride = (Y * Z, Z, 1)[axis]
        f = self._flat(idx) + shift * stride
        if len(f) and (f.min() < 0 or f.max() >= X * Y * Z):
            raise ValueError("a solid node's neighbour falls outside the array")
        return self.bitmask.reshape(-1)[f]

    def finalize(self) -> None:
        """Solid boundary / interior solids, fluid neighbours and tables
        (``immersedBoundary.py:54-152``)."""
        solid = np.stack(np.nonzero(self.bitmask), axis=-1).astype(np.int32)
        elig = self._eligible()
        is_bd = np.zeros(len(solid), dtype=bool)
        for a in elig:
            is_bd |= self._neighbour_bits(solid, a, -1) == 0
            is_bd |= self._neighbour_bits(solid, a, +1) == 0
        self.solid_indices = solid
        self.solid_boundary = solid[is_bd]
        self.interior_solids = solid[~is_bd]
        # fluid neighbours of each boundary node, (node, eligible axis, side)
        nb = np.zeros((len(self.solid_boundary), len(elig), 2), dtype=np.int8)
        fluid = []
        for d, a in enumerate(elig):
            for i, s in enumerate((-1, 1)):
                bits = self._neighbour_bits(self.solid_boundary, a, s)
                nb[:, d, i] = bits
                f = self.solid_boundary[bits == FLUID_CELL].copy()
                f[:, a] += s          # grid index (may be -1 / P on the outer layer)
                fluid.append(f)
        self.fluid_neighbors = nb
        allf = np.concatenate(fluid, axis=0) if fluid else np.zeros((0, 3), np.int32)
        self.fluid_boundary = np.unique(allf, axis=0)
        self.define_boundary_value_and_type_arrays(self.solid_boundary)
        self._ops_dev = None

    # -------------------------------------------------------------- tables
    def resolved_ops(self) -> tuple:
        """(pde_ib_op records, interior-solid flat node indices)."""
        if not hasattr(self, "solid_boundary"):
            raise AssertionError("call finalize() before applying the boundary")
        X, Y, Z = self.grid_shape
        C = self.inputs[0]
        sb = self.solid_boundary.astype(np.int64)
        nodes = (sb[:, 0] * Y + sb[:, 1]) * Z + sb[:, 2]
        mask = np.zeros(len(sb), dtype=np.int32)
        for d in range(self.fluid_neighbors.shape[1]):
            for i in range(2):
                mask |= (self.fluid_neighbors[:, d, i] == FLUID_CELL).astype(np.int32) << (2 * d + i)
        ops = np.zeros(len(sb) * C, N.IB_OP_DTYPE)
        ops["node"] = np.repeat(nodes, C)
        ops["comp"] = np.tile(np.arange(C, dtype=np.int32), len(sb))
        ops["kind"] = self.boundary_type.reshape(-1).astype(np.int32)
        ops["value"] = self.boundary_value.reshape(-1)
        ops["mask"] = np.repeat(mask, C)
        isd = self.interior_solids.astype(np.int64)
        fill = (isd[:, 0] * Y + isd[:, 1]) * Z + isd[:, 2]
        return ops, fill

    def _device_tables(self, device):
        if self._ops_dev is None or self._ops_dev[0].device != device:
            ops, fill = self.resolved_ops()
            self._ops_dev = (torch.from_numpy(ops.view(np.uint8)).to(device),
                             torch.from_numpy(fill).to(device), len(ops), len(fill))
        return self._ops_dev

    def set_BC(self, face_ids, value, boundary_type, outputs_ids) -> None:
        super().set_BC(face_ids, value, boundary_type, outputs_ids)
        self._ops_dev = None

    # -------------------------------------------------------------- apply
    def _launch(self, x: torch.Tensor, out: torch.Tensor, fill_value: float) -> None:
        ops, fill, nops, nfill = self._device_tables(x.device)
        elig = self._eligible()
        X, Y, Z = self.grid_shape
        node_stride = (Y * Z, Z, 1)
        st = [node_stride[a] for a in elig] + [0] * (3 - len(elig))
        N.call("pde_immersed_apply_f32", N.dev_ptr(x), N.dev_ptr(out), ops.data_ptr(), nops,
               fill.data_ptr(), nfill, np.float32(fill_value), self.inputs[0], len(elig),
               st[0], st[1], st[2], np.float32(self.dx), N.stream_of(x))

    @setup
    def initialize_kernel(self, input_array: torch.Tensor, *args, **kwargs) -> None:
        self._check_input(input_array)
        if tuple(input_array.shape[:3]) != self.grid_shape:
            raise AssertionError("field shape differs from the one the boundary was built for")
        self._device_tables(input_array.device)

    @before_forward
    def copy_array(self, input_array: torch.Tensor, *args, **kwargs) -> None:
        N.check_field(input_array, "input_array")
        N.call("pde_copy_f32", N.dev_ptr(input_array), N.dev_ptr(self.output_array),
               input_array.numel(), N.stream_of(input_array))

    def forward(self, input_array: torch.Tensor, fill_value: float = 0.0, *args, **kwargs):
        self._launch(input_array, self.output_array, fill_value)
        return self.output_array

    def apply_inplace(self, u: torch.Tensor, fill_value: float = 0.0) -> torch.Tensor:
        """The boundary applied to u itself (one launch, no copy)."""
        N.check_field(u, "u")
        self._check_input(u)
        self._launch(u, u, fill_value)
        return u
