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
from ..FDM.boundary.boundary_functions import FunctionValue
from ..FDM.boundary.device_table import DeviceBoundaryTable
from ..FDM.boundary.function_values import FunctionGroup, FunctionSlots
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

    def _slot_bases(self):
        bases, n = {}, 0
        for key, fb in self.func_groups.items():
            if isinstance(fb.func, FunctionValue):
                bases[key] = n
                n += len(fb.face_ids)
        return bases

    def resolved_ops(self) -> np.ndarray:
        bases = self._slot_bases()
        fgs = [(fb.face_ids, fb.boundary_type, fb.varID, fb.func, bases.get(key))
               for key, fb in self.func_groups.items()]
        return legacy_ops(self.grid_shape, self.inputs[0], self.boundary_ijk_indices,
                          self.boundary_interior, self.boundary_type, self.boundary_value,
                          np.float32(self.dx), fgs)

    def function_groups(self) -> list[FunctionGroup]:
        """Function groups given by Python callables: their nodes (the boundary
        node list's, shifted groups included) and slots."""
        return [FunctionGroup(key, self.func_groups[key].func.func,
                              self.boundary_ijk_indices[self.func_groups[key].face_ids], base, 1,
                              self.func_groups[key].varID)
                for key, base in self._slot_bases().items()]

    def _build_table(self, device) -> None:
        self.table = DeviceBoundaryTable(self.resolved_ops(), device)
        fg = self.function_groups()
        self._fslots = FunctionSlots(fg, device, np.float32(self.dx), self.grid_coordinates) \
            if fg else None

    def __call__(self, input_array: torch.Tensor, t: float = 0.0, params=None):
        if params is None:
            params = {}
        return super().__call__(input_array, t, params)

    @setup
    def to_device(self, input_array, *args, **kwargs) -> None:
        self._check_input(input_array)
        if tuple(input_array.shape[:3]) != self.grid_shape:
            raise AssertionError("field shape differs from the one the boundary was built for")
        self._set_time(input_array, None)
        self._build_table(input_array.device)

    @before_forward
    def set_default_params(self, input_array, t, params, **kwargs) -> None:
        """Reference ``gridBoundary.py:182-191``: default params per function
        group, then the time (a number, or a 1-element float32 device array)."""
        for key in self.func_groups:
            params.setdefault(key, 0)
        self._params = params
        if self.table is None:
            self._build_table(input_array.device)
        self._set_time(input_array, t)

    def _set_time(self, u, t) -> None:
        """t: None keeps the current time array; a tensor (the caller's
        1-element float32 device array) becomes it; a number is written into
        the boundary's own array (never into a caller's)."""
        own = getattr(self, "_t_own", None)
        if own is None or own.device != u.device:
            own = self._t_own = torch.zeros(1, dtype=torch.float32, device=u.device)
            if not hasattr(self, "t") or self.t.device != u.device:
                self.t = own
        if isinstance(t, torch.Tensor):
            if t.numel() != 1 or t.dtype != torch.float32 or t.device != u.device:
                raise AssertionError("t must be a 1-element float32 array on the field's device")
            self.t = t
        elif t is not None:
            own.fill_(float(np.float32(t)))
            self.t = own

    def _values(self, u: torch.Tensor):
        """None for a table of constants (the kernels need no time), else the
        time array, followed by the function groups' values on u."""
        if self._fslots is not None:
            return self._fslots.evaluate(u, self.t, getattr(self, "_params", None))
        return self.t if self.table.needs_t else None

    def forward(self, input_array, t=0.0, params=None, **kwargs):
        return self.table.apply(input_array, self.output_array, self._values(input_array))

    def _ensure_table(self, u: torch.Tensor) -> None:
        self._check_input(u)
        if self.table is None or self.table.device != u.device:
            self._build_table(u.device)

    def _time_tensor(self, u: torch.Tensor, t=None):
        """The device array the boundary values read (None when every value is
        a constant)."""
        self._ensure_table(u)
        self._set_time(u, t)
        return self._values(u)

    def apply_inplace(self, u: torch.Tensor, t=None) -> torch.Tensor:
        """Apply the boundary to u itself (no copy), as the fused steps do
        (t: None keeps the time of the last call)."""
        self._check_input(u)
        tt = self._time_tensor(u, t)          # builds the table when groups changed
        return self.table.apply_inplace(u, tt)
