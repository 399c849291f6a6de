"""Functional GridBoundary (reference
``pm/FDM/boundary/functional_boundary.py:57-294``).

Groups: '-X', '+X', ... are the face layers (index g or P-1-g) with the other
eligible axes restricted to [g, P-1-g]; 'ALL' is their union (``:111-137``).
``set_BC`` appends a group; ``__call__`` returns ``copy(x)`` with the groups
applied in insertion order.  Nodes in no group are untouched (there is no
default).  Dirichlet writes the whole node vector with the selected components
replaced and fills the first ghost layer with ``2 v - x[adj]``; Neumann fills it
with ``-sign(s) 2 dx v + x[adj]`` where ``v`` is the derivative along +axis.

Group names are ``"<face>_<TYPE>"`` (made unique); the reference builds them
from ``None`` by mistake (``:168``).  The names only key the ``params`` dict.
"""
from __future__ import annotations

from typing import Any

import numpy as np
import torch

from ...stencil import before_forward, setup
from ...utils import get_unique_key
from ..module.explicitUniformGridStencil import ExplicitUniformGridStencil
from .boundary_functions import FunctionValue, as_boundary_value
from .device_table import DeviceBoundaryTable
from .flags import BC_SRING, DIRICHLET, VON_NEUMANN
from .function_values import FunctionGroup, FunctionSlots
from .tables import functional_ops


class FunctionBC:
    """One boundary group (reference ``functional_boundary.py:25-55``)."""

    def __init__(self, face_ids, boundary_type, output_ids, func):
        self.face_ids = np.asarray(face_ids, dtype=np.int64).ravel()
        self.boundary_type = int(boundary_type)
        self.output_ids = output_ids
        self.func = func


class GridBoundary(ExplicitUniformGridStencil):
    """Ghost-cell Dirichlet / Neumann boundaries around a padded vec(C) field."""

    def __init__(self, field: torch.Tensor, dx: float, ghost_cells: int,
                 grid_coordinates=None) -> None:
        if field.dim() != 4:
            raise AssertionError("GridBoundary applies to vector fields (X, Y, Z, C)")
        inputs = int(field.shape[3])
        self.dimension = sum(1 for s in field.shape[:3] if s > 1)
        super().__init__(inputs, inputs, dx, ghost_cells)
        self.grid_coordinates = grid_coordinates
        self.grid_shape = tuple(int(s) for s in field.shape[:3])
        self.nodeIDs = np.arange(int(np.prod(self.grid_shape)), dtype=np.int32)
        self.groups: dict[str, np.ndarray] = {}
        self.func_groups: dict[str, FunctionBC] = {}
        self.table: DeviceBoundaryTable | None = None
        self.define_exterior_groups()

    def define_exterior_groups(self) -> None:
        g = self.ghost_cells
        flat = self.nodeIDs.reshape(self.grid_shape)
        parts = []
        for i, axis in enumerate("XYZ"):
            if self.grid_shape[i] <= 1:
                continue
            for prefix, sign in (("-", -1), ("+", 1)):
                sl = [slice(None)] * 3
                sl[i] = g if sign == -1 else -(g + 1)
                for x, p in enumerate(self.grid_shape):
                    if p > 1 and x != i:
                        sl[x] = slice(g, -g) if g > 0 else slice(None)
                self.groups[prefix + axis] = flat[tuple(sl)].ravel()
                parts.append(self.groups[prefix + axis])
        self.groups["ALL"] = np.unique(np.concatenate(parts)).astype(np.int32)

    def set_BC(self, face_ids, value, boundary_type: int, output_ids=None) -> None:
        if boundary_type not in (DIRICHLET, VON_NEUMANN):
            raise ValueError(f"unknown boundary type {boundary_type}")
        if output_ids is None:
            output_ids = tuple(range(self.inputs[0]))
        if not isinstance(output_ids, (int, list, tuple)):
            raise AssertionError("output id must be an integer")
        ids_list = [output_ids] if isinstance(output_ids, int) else list(output_ids)
        if not all(0 <= int(c) < self.inputs[0] for c in ids_list):
            raise AssertionError("output ids out of range")
        func = as_boundary_value(value)
        if isinstance(face_ids, str):
            if face_ids not in self.groups:
                raise AssertionError(f"unknown group {face_ids!r}")
            name = get_unique_key(self.func_groups, f"{face_ids}_{BC_SRING[boundary_type]}")
            face_ids = self.groups[face_ids]
        else:
            if not isinstance(face_ids, (np.ndarray, list, tuple, int)):
                raise AssertionError("face ids must be a group name or node ids")
            name = get_unique_key(self.func_groups, "BC_Group")
            face_ids = np.atleast_1d(np.array(face_ids, dtype=np.int64))
            if face_ids.ndim > 1:
                raise AssertionError("face_ids should be 1D")
            if face_ids.size and (face_ids.min() < 0 or face_ids.max() >= len(self.nodeIDs)):
                raise IndexError("face ids must be flat node ids inside the padded grid")
        self.func_groups[name] = FunctionBC(face_ids, boundary_type, output_ids, func)
        self.table = None

    def dirichlet_BC(self, face_ids, value, output_ids=None) -> None:
        self.set_BC(face_ids, value, DIRICHLET, output_ids)

    def vonNeumann_BC(self, face_ids, value, output_ids=None) -> None:
        self.set_BC(face_ids, value, VON_NEUMANN, output_ids)

    def no_slip(self, face_ids, output_ids=None) -> None:
        self.set_BC(face_ids, 0.0, DIRICHLET, output_ids)

    def impermeable(self, group: str) -> None:
        if self.inputs[0] != self.dimension:
            raise AssertionError(
                "impermeable needs a velocity field: one component per grid dimension")
        if group not in {"-X", "+X", "-Y", "+Y", "-Z", "+Z"}:
            raise AssertionError("impermeable takes one face group: -X, +X, -Y, +Y, -Z or +Z")
        axis = "XYZ".index(group[-1])
        self.vonNeumann_BC(group, 0.0)
        self.dirichlet_BC(group, 0.0, axis)

    def _slot_bases(self):
        """First value slot of each function group (FunctionValue), in order."""
        bases, n = {}, 0
        for key, fb in self.func_groups.items():
            if isinstance(fb.func, FunctionValue):
                bases[key] = n
                n += len(fb.face_ids) * self.inputs[0]
        return bases

    def resolved_ops(self) -> np.ndarray:
        """The host-resolved op table (struct pde_bc_op records)."""
        bases = self._slot_bases()
        groups = [(fb.face_ids, fb.boundary_type, fb.output_ids, fb.func, bases.get(key))
                  for key, fb in self.func_groups.items()]
        return functional_ops(self.grid_shape, self.inputs[0], self.ghost_cells, groups,
                              np.float32(self.dx))

    def function_groups(self) -> list[FunctionGroup]:
        """The groups whose values a Python function gives, with their node
        indices and slots."""
        Y, Z = self.grid_shape[1], self.grid_shape[2]
        out = []
        for key, base in self._slot_bases().items():
            fb = self.func_groups[key]
            ids = fb.face_ids
            ijk = np.stack([ids // (Y * Z), (ids // Z) % Y, ids % Z], axis=1)
            out.append(FunctionGroup(key, fb.func.func, ijk, base, self.inputs[0], None))
        return out

    def _build_table(self, device) -> None:
        self.table = DeviceBoundaryTable(self.resolved_ops(), device)
        fg = self.function_groups()
        self._fslots = FunctionSlots(fg, device, np.float32(self.dx), self.grid_coordinates) \
            if fg else None

    def __call__(self, input_array: torch.Tensor, t=None, params: dict[str, Any] | None = None):
        if params is None:
            params = {}
        return super().__call__(input_array, t, params)

    @setup(order=1)
    def initialize_kernel(self, input_array: torch.Tensor, *args, **kwargs) -> None:
        self._check_input(input_array)
        if tuple(input_array.shape[:3]) != self.grid_shape:
            raise AssertionError("field shape differs from the one the boundary was built for")
        self._t_own = torch.zeros(1, dtype=torch.float32, device=input_array.device)
        self.t = self._t_own
        self._build_table(input_array.device)

    @staticmethod
    def _check_time_array(t: torch.Tensor, u: torch.Tensor) -> None:
        """A caller's time array is read on the device by the boundary kernels:
        it must be one float32 value on the field's device."""
        if t.numel() != 1 or t.dtype != torch.float32 or t.device != u.device:
            raise AssertionError("t must be a 1-element float32 array on the field's device")

    @before_forward
    def set_default_params(self, input_array, t, params, **kwargs) -> None:
        for key in self.func_groups:
            params.setdefault(key, 0)
        # the params of the last call also feed the fused steps' evaluations
        self._params = params
        if self.table is None:  # groups changed after setup
            self._build_table(input_array.device)
        self._set_time(input_array, t)

    def _set_time(self, u: torch.Tensor, t) -> None:
        """t: None keeps the current time array; a tensor (the caller's
        1-element float32 device array) becomes it; a number is written into
        the boundary's own array (never into a caller's)."""
        own = getattr(self, "_t_own", None)
        if own is None or own.device != u.device:
            own = self._t_own = torch.zeros(1, dtype=torch.float32, device=u.device)
            if not hasattr(self, "t") or self.t.device != u.device:
                self.t = own
        if isinstance(t, torch.Tensor):
            self._check_time_array(t, u)
            self.t = t
        elif t is not None:
            own.fill_(float(np.float32(t)))
            self.t = own

    def _values(self, u: torch.Tensor) -> torch.Tensor:
        """The array the kernels read: the time, then (when some group's value
        is a Python function) every function group's values on u."""
        fs = getattr(self, "_fslots", None)
        if fs is None:
            return self.t
        return fs.evaluate(u, self.t, getattr(self, "_params", None))

    def forward(self, input_array, t=None, params=None, **kwargs):
        return self.table.apply(input_array, self.output_array, self._values(input_array))

    def _ensure_table(self, u: torch.Tensor) -> None:
        self._check_input(u)
        if self.table is None or self.table.device != u.device:
            self._build_table(u.device)

    def _time_tensor(self, u: torch.Tensor, t=None):
        """The device array the boundary values read (time, then function
        values evaluated on u, the field before the boundary)."""
        self._set_time(u, t)
        self._ensure_table(u)
        return self._values(u)

    def apply_inplace(self, u: torch.Tensor, t=None) -> torch.Tensor:
        """Apply the boundary to ``u`` itself (no copy): the form the fused time
        steps use.  Same values as ``__call__`` on every written scalar."""
        self._check_input(u)
        self.set_default_params(u, t, getattr(self, "_params", None) or {})
        if self.table.device != u.device:
            self._build_table(u.device)
        return self.table.apply_inplace(u, self._values(u))
