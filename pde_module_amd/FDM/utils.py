"""Grid helpers (reference ``pm/FDM/utils.py:6-27``)."""
from __future__ import annotations


def eligible_dims_and_shift(grid_shape, ghost_cells):
    """Axes with more than one point, ascending, and the ghost shift per axis
    (g on eligible axes, 0 elsewhere).  ``(3, 5, 1)`` with g > 0 gives
    ``((0, 1), (g, g, 0))``."""
    dims = tuple(i for i, x in enumerate(grid_shape) if x > 1)
    return dims, tuple(ghost_cells if x > 1 else 0 for x in grid_shape)


def grid_shape_with_no_ghost_cells(grid_shape, ghost_cells):
    return tuple(a - 2 * ghost_cells if a > 1 else a for a in grid_shape)


def get_ghost_shape(grid_shape, ghost_cells):
    return tuple(a + 2 * ghost_cells if a > 1 else a for a in grid_shape)


def dimension_of(grid_shape):
    """Number of grid axes with more than one point.  Only the first three
    entries are grid axes: a torch field's ``shape`` carries its element dims
    after them (a Warp array's shape stops at the grid), so ``Grad(2, u.shape,
    ...)`` means the same here as in the reference."""
    return sum(1 for s in tuple(grid_shape)[:3] if s > 1)
