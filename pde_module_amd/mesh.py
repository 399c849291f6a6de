"""Uniform grid coordinates and padded field allocation (the layout contract of
reference ``pm/mesh/uniformGridMesh/grid.py:7-108``,
``pm/mesh/uniformGridMesh/functions.py:25-90`` and
``pm/mesh/mesh_functions.py:108-159``).

Node coordinates are ``x_i = i * dx - origin`` (FP32), extended by ``k * dx`` on
both sides for the ghost layers; a field is a tensor of padded shape
``(X, Y, Z)`` + element shape, with initial values evaluated on the full
meshgrid, ghosts included.  Unstructured-mesh topology (cells, faces,
connectivity) is not part of the finite-difference path and is not built.
"""
from __future__ import annotations

from math import prod
from typing import Callable, Optional

import numpy as np
import torch


def _ghost_coords(dx, vectors, g):
    if g == 0:
        return list(vectors)
    out = []
    for v in vectors:
        if len(v) == 1:
            out.append(v)
            continue
        left = np.array([v[0] - n * dx for n in range(1, g + 1)], dtype=v.dtype)
        right = np.array([v[-1] + n * dx for n in range(1, g + 1)], dtype=v.dtype)
        out.append(np.concatenate((left, v, right)).astype(v.dtype))
    return out


class UniformGridMesh:
    def __init__(self, dx: float, nodes_per_axis, origin=None, ghost_cells: Optional[int] = None,
                 float_dtype=np.float32) -> None:
        dx = float_dtype(dx)
        self.dx = dx
        self.area = dx ** 2
        self.volume = dx ** 3
        g = 0 if ghost_cells is None else ghost_cells
        if not (isinstance(g, int) and g >= 0):
            raise AssertionError("ghost_cells must be a non-negative int")
        self.ghost_cells = g
        self.float_dtype = float_dtype
        self.origin = (np.zeros(3, dtype=float_dtype) if origin is None
                       else np.array(origin, dtype=float_dtype))
        if len(self.origin) != 3 or len(nodes_per_axis) != 3:
            raise AssertionError("nodes_per_axis and origin are 3-tuples")
        base = tuple(np.arange(0, n, dtype=float_dtype) * dx - o
                     for n, o in zip(nodes_per_axis, self.origin))
        self.coordinate_vectors = _ghost_coords(dx, base, g)
        mg = np.meshgrid(*self.coordinate_vectors, indexing="ij")
        self.nodal_grid = np.stack(mg, axis=-1)
        self.meshgrid = [self.nodal_grid[..., i] for i in range(3)]
        self.nodes_per_axis = tuple(len(c) for c in self.coordinate_vectors)
        self.grid_shape = self.nodes_per_axis
        self.dimension = sum(1 for n in self.nodes_per_axis if n > 1)
        if not 1 <= self.dimension <= 3:
            raise AssertionError("grid must have 1 to 3 axes with more than one node")
        self.num_nodes = prod(self.nodes_per_axis)
        self.num_cells = prod(n - 1 for n in self.nodes_per_axis if n > 1)


def create_structured_field(mesh: UniformGridMesh, type: str = "node", num_outputs=1,
                            AoS: bool = True, func: Optional[Callable] = None,
                            device="cuda", **kwargs) -> torch.Tensor:
    """A padded FP32 field on ``device``: zeros, or ``func(*mesh.meshgrid)``."""
    if type == "node":
        shape = tuple(mesh.nodal_grid.shape[:-1])
    elif type == "cell":
        shape = tuple(n - 1 for n in mesh.nodal_grid.shape[:-1])
    else:
        raise ValueError("Valid types are: node and cell")
    if not AoS:
        raise NotImplementedError("fields are array-of-structures only")
    elem = (num_outputs,) if isinstance(num_outputs, int) else tuple(num_outputs)
    if len(elem) not in (1, 2):
        raise ValueError("num_outputs is an int (vector) or a 2-tuple (matrix)")
    full = shape + elem
    if func is None:
        return torch.zeros(full, dtype=torch.float32, device=device)
    arr = np.asarray(func(*mesh.meshgrid, **kwargs), dtype=np.float32).reshape(full)
    return torch.from_numpy(np.ascontiguousarray(arr)).to(device)


# name kept for code written against the reference
create_structured_warp_field = create_structured_field
