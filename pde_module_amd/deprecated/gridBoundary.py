"""Legacy GridBoundary (reference ``pm/deprecated/gridBoundary.py:11-304``), still
the boundary of the reference's heat and standing-wave examples.

Boundary nodes are every non-ghost node with an eligible index at g or
P-1-g, listed once in lexicographic (i, j, k) order; groups '-X', '+X', ... and
'ALL' are index sets into that list; each node's interior direction is +1 on the
axes of its '-' faces and -1 on those of its '+' faces.  Every (node,
component) entry defaults to Dirichlet 0.  ``__call__`` returns copy(x) with:
Dirichlet ``out[b][c] = v`` and ``out[b - s e_a][c] = 2 v - x[b + s e_a][c]``;
Neumann ``out[b - s e_a][c] = -sign(s) 2 dx v + x[b + s e_a][c]``, for each axis
with s != 0.
"""
from __future__ import annotations

import numpy as np
import torch

from ..FDM.boundary.boundary import Boundary
from ..FDM.boundary.device_table import DeviceBoundaryTable
from ..FDM.boundary.tables import legacy_ops
from ..stencil import before_forward, setup


class GridBoundary(Boundary):
    def __init__(self, field: torch.Tensor, dx: float, ghost_cells: int,
                 grid_coordinates=None) -> None:
        super().__init__(field, dx, ghost_cells)
        self.grid_coordinates = grid_coordinates
        self.table = None
        self.define_boundary_ijk_indices()
        self.define_groups()
        self.define_interior_adjacency()
        self.define_boundary_value_and_type_arrays(self.boundary_ijk_indices)

    def define_boundary_ijk_indices(self, *args, **kwargs) -> None:
        ng = self.grid_shape_without_ghost
        parts = []
        for i in range(3):
            if ng[i] == 1:
                continue
            for lim in (0, ng[i] - 1):
                side = list(ng)
                side[i] = 1
                idx = np.moveaxis(np.indices(side, dtype=np.int32), 0, -1).reshape(-1, 3)
                idx[:, i] = lim
                for ax in range(3):
                    if ng[ax] != 1:
                        idx[:, ax] += self.ghost_cells
                parts.append(idx)
        self.boundary_ijk_indices = np.unique(np.concatenate(parts, axis=0), axis=0).astype(np.int32)

    def define_groups(self, *args, **kwargs) -> None:
        self.boundary_ids = np.arange(len(self.boundary_ijk_indices), dtype=np.int32)
        self.groups["ALL"] = self.boundary_ids
        ng = self.grid_shape_without_ghost
        for axis, name in enumerate("XYZ"):
            if ng[axis] == 1:
                continue
            lims = (self.ghost_cells, ng[axis] - 1 + self.ghost_cells)
            for lim, parity in zip(lims, "-+"):
                self.groups[parity + name] = self.boundary_ids[
                    self.boundary_ijk_indices[:, axis] == lim]

    def define_interior_adjacency(self) -> None:
        self.boundary_interior = np.zeros((len(self.boundary_ijk_indices), 3), dtype=np.int32)
        for i, name in enumerate("XYZ"):
            if self.grid_shape_without_ghost[i] == 1:
                continue
            for parity, sign in (("-", 1), ("+", -1)):
                self.boundary_interior[self.groups[parity + name], i] = sign

    def shift_group(self, group: str, shift: int, axis: int) -> None:
        """Move a group's nodes along an axis, keeping their interior directions
        (reference ``gridBoundary.py:118-130``)."""
        if axis >= self.dimension or group not in self.groups:
            raise AssertionError("bad group or axis")
        self.boundary_ijk_indices[self.groups[group], axis] += shift
        self.table = None

    def resolved_ops(self) -> np.ndarray:
        return legacy_ops(self.grid_shape, self.inputs[0], self.boundary_ijk_indices,
                          self.boundary_interior, self.boundary_type, self.boundary_value,
                          np.float32(self.dx))

    def __call__(self, input_array: torch.Tensor, t: float = 0.0, params=None):
        if params is None:
            params = {}
        return super().__call__(input_array, t, params)

    @setup
    def to_device(self, input_array, *args, **kwargs) -> None:
        self._check_input(input_array)
        if tuple(input_array.shape[:3]) != self.grid_shape:
            raise AssertionError("field shape differs from the one the boundary was built for")
        self.t = torch.zeros(1, dtype=torch.float32, device=input_array.device)
        self.table = DeviceBoundaryTable(self.resolved_ops(), input_array.device)

    @before_forward
    def set_default_params(self, input_array, t, params, **kwargs) -> None:
        if self.table is None:
            self.table = DeviceBoundaryTable(self.resolved_ops(), input_array.device)

    def forward(self, input_array, t=0.0, params=None, **kwargs):
        return self.table.apply(input_array, self.output_array, None)

    def _ensure_table(self, u: torch.Tensor) -> None:
        self._check_input(u)
        if self.table is None or self.table.device != u.device:
            self.table = DeviceBoundaryTable(self.resolved_ops(), u.device)

    def _time_tensor(self, u: torch.Tensor, t=None):
        return None  # legacy tables hold constants only

    def apply_inplace(self, u: torch.Tensor, t=0.0) -> torch.Tensor:
        """Apply the boundary to u itself (no copy), as the fused steps do."""
        self._check_input(u)
        if self.table is None:
            self.table = DeviceBoundaryTable(self.resolved_ops(), u.device)
        return self.table.apply_inplace(u, None)
