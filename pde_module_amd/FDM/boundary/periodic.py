"""Periodic ghost fill (no counterpart in the reference, which has no periodic
boundary).  Used for periodic grids such as the 3D operator benchmark."""
from __future__ import annotations

import torch

from ... import _native as N


class PeriodicBoundary:
    """``PeriodicBoundary(ghost_cells)(u)`` fills u's ghost layers in place from
    the opposite interior layers, axis by axis (edges and corners included),
    and returns u."""

    def __init__(self, ghost_cells: int) -> None:
        self.ghost_cells = ghost_cells

    def __call__(self, u: torch.Tensor) -> torch.Tensor:
        N.check_field(u, "field")
        if u.dim() < 4:
            raise ValueError("fields are (X, Y, Z, C...) tensors")
        X, Y, Z = (int(s) for s in u.shape[:3])
        comps = int(u[0, 0, 0].numel())
        N.call("pde_periodic_fill_f32", N.dev_ptr(u), X, Y, Z, comps, self.ghost_cells,
               N.stream_of(u))
        return u
