"""Table-driven boundary base used by the legacy GridBoundary (reference
``pm/FDM/boundary/boundary.py:11-222``).

Every boundary (node, component) entry has a type and a value.  Types start as
DIRICHLET and values as 0 (``boundary.py:59-65``), so an entry never set is a
Dirichlet-0 boundary.  ``set_BC`` overwrites entries of a group or of explicit
entry ids.
"""
from __future__ import annotations

import numpy as np
import torch

from ...utils import get_unique_key
from ..module.explicitUniformGridStencil import ExplicitUniformGridStencil
from .boundary_functions import ConstantValue, as_boundary_value
from .flags import Boundary_Types


class FunctionBC:
    """One function-valued group of the legacy boundary (reference
    ``pm/deprecated/func_boundary.py:47-81``): entry ids into the boundary
    node list, the type, the component it writes and the value."""

    def __init__(self, face_ids, boundary_type, varID, func):
        self.face_ids = np.atleast_1d(np.asarray(face_ids, dtype=np.int64)).ravel()
        self.boundary_type = int(boundary_type)
        self.varID = int(varID)
        self.func = func


class Boundary(ExplicitUniformGridStencil):
    boundary_type: np.ndarray
    boundary_value: np.ndarray

    def __init__(self, field: torch.Tensor, dx: float, ghost_cells: int) -> None:
        if field.dim() != 4:
            raise AssertionError("input must be vector type")
        inputs = int(field.shape[3])
        super().__init__(inputs, inputs, dx)
        if not (type(ghost_cells) is int and ghost_cells > 0):
            raise AssertionError("ghost_cells must be a positive int")
        self.ghost_cells = ghost_cells
        self.grid_shape = tuple(int(s) for s in field.shape[:3])
        self.grid_shape_without_ghost = self.grid_shape_with_no_ghost_cells(self.grid_shape,
                                                                            ghost_cells)
        self.dimension = self.calculate_dimension_from_grid_shape(self.grid_shape)
        self.groups: dict[str, np.ndarray] = {}
        self.func_groups: dict = {}

    def define_boundary_value_and_type_arrays(self, indices: np.ndarray) -> None:
        shape = (len(indices),) + self.input_dtype_shape
        self.boundary_ids = np.arange(len(indices))
        self.boundary_value = np.zeros(shape, dtype=np.float32)
        self.boundary_type = np.ones_like(self.boundary_value, dtype=np.int8)
        self.groups["ALL"] = self.boundary_ids

    def _check_output_ids(self, output_ids):
        if output_ids is None:
            return slice(None)
        if isinstance(output_ids, (int, np.integer)):
            if not 0 <= output_ids < self.inputs[0]:
                raise AssertionError("output id out of range")
            return int(output_ids)
        if isinstance(output_ids, (list, tuple, np.ndarray)):
            output_ids = np.array(output_ids, dtype=np.int32)
            if not np.all((0 <= output_ids) & (output_ids < self.inputs[0])):
                raise AssertionError("output ids out of range")
            return output_ids
        raise TypeError(
            f"output ids are an int, a sequence of ints or None, not {type(output_ids).__name__}")

    def set_BC(self, face_ids, value: float, boundary_type: int, outputs_ids) -> None:
        if isinstance(face_ids, str):
            if face_ids not in self.groups:
                raise AssertionError(f"unknown group {face_ids!r}")
            face_ids = self.groups[face_ids]
        else:
            if not isinstance(face_ids, (np.ndarray, list, tuple, int)):
                raise AssertionError("face ids must be a group name or entry ids")
            face_ids = np.array(face_ids, dtype=np.int32)
        outputs_ids = self._check_output_ids(outputs_ids)
        if not isinstance(value, float):
            raise AssertionError("legacy boundary values must be float")
        # NumPy indexing exactly as the reference (``boundary.py:111-113``): an
        # id array with a component array pairs them element-wise.
        self.boundary_type[face_ids, outputs_ids] = boundary_type
        self.boundary_value[face_ids, outputs_ids] = value
        self.table = None

    def set_func_BC(self, face_ids, func, boundary_type, outputs_id) -> None:
        """A function-valued group (reference ``boundary.py:115-143``): applied
        after the table, groups in the order set, one component each.  The
        group is keyed (for ``params``) by its name, or a fresh "BC_Group"
        name for explicit entry ids; setting a function on the same group name
        again replaces it.  ``func`` is a Python callable
        ``func(current_values, nodeID, varID, coordinates, t, dx, params)``
        evaluated on the device for the whole group, or a value record
        (get_ramp_func, get_sine_func)."""
        if type(outputs_id) is not int:
            raise AssertionError("a boundary function writes one component: give an int id")
        if not 0 <= outputs_id < self.inputs[0]:
            raise AssertionError("output id out of range")
        if boundary_type not in (Boundary_Types.DIRICHLET, Boundary_Types.VON_NEUMANN):
            raise ValueError(f"unknown boundary type {boundary_type}")
        value = as_boundary_value(func)
        if isinstance(face_ids, str):
            if face_ids not in self.groups:
                raise AssertionError(f"unknown group {face_ids!r}")
            name = face_ids
            face_ids = self.groups[face_ids]
        else:
            if not isinstance(face_ids, (np.ndarray, list, tuple, int)):
                raise AssertionError("face ids must be a group name or entry ids")
            name = get_unique_key(self.groups, "BC_Group")
            face_ids = np.atleast_1d(np.array(face_ids, dtype=np.int64))
            if face_ids.size and (face_ids.min() < 0 or face_ids.max() >= len(self.boundary_ids)):
                raise IndexError("entry ids must index the boundary node list")
        self.func_groups[name] = FunctionBC(face_ids, boundary_type, outputs_id, value)
        self.table = None

    def dirichlet_BC(self, group, value, outputs_ids=None) -> None:
        if isinstance(value, (float, int)) and not isinstance(value, bool):
            self.set_BC(group, float(value), Boundary_Types.DIRICHLET, outputs_ids)
        elif isinstance(value, ConstantValue):
            self.set_BC(group, float(value.constant), Boundary_Types.DIRICHLET, outputs_ids)
        else:
            self.set_func_BC(group, value, Boundary_Types.DIRICHLET, outputs_ids)

    def vonNeumann_BC(self, group, value, outputs_ids=None) -> None:
        """Numbers go to the table as in the reference; a function or value
        record (which the reference's legacy vonNeumann_BC refuses) becomes a
        Neumann function group."""
        if isinstance(value, (float, int)) and not isinstance(value, bool):
            self.set_BC(group, float(value), Boundary_Types.VON_NEUMANN, outputs_ids)
        elif isinstance(value, ConstantValue):
            self.set_BC(group, float(value.constant), Boundary_Types.VON_NEUMANN, outputs_ids)
        else:
            self.set_func_BC(group, value, Boundary_Types.VON_NEUMANN, outputs_ids)

    def no_slip(self, group) -> None:
        if self.inputs[0] != self.dimension:
            raise AssertionError(
                "this needs a velocity field: one component per grid dimension")
        self.dirichlet_BC(group, 0.0)

    def impermeable(self, group: str) -> None:
        if self.inputs[0] != self.dimension:
            raise AssertionError(
                "this needs a velocity field: one component per grid dimension")
        if group not in self.groups or group not in {"-X", "+X", "-Y", "+Y", "-Z", "+Z"}:
            raise AssertionError("impermeable takes one face group: -X, +X, -Y, +Y, -Z or +Z")
        axis = "XYZ".index(group[-1])
        self.vonNeumann_BC(group, 0.0)
        self.dirichlet_BC(group, 0.0, axis)

    def check_boundary_types(self) -> bool:
        return not np.any(self.boundary_type == Boundary_Types.NO_BC)


__all__ = ["Boundary", "FunctionBC", "get_unique_key"]
