"""Viscous damping layer (reference ``pm/FDM/module/viscousDampingLayer.py:13-161``)."""
from __future__ import annotations

import numpy as np

from ...stencil import setup
from ..functional import viscousDampingLayer
from ..kernel import create_dampingLayer_kernel
from ..layers import layer_groups, layer_points, point_table
from .explicitUniformGridStencil import ExplicitUniformGridStencil


class ViscousDampingLayer(ExplicitUniformGridStencil):
    """``ViscousDampingLayer(inputs, num_layers, grid_shape, dx, ghost_cells, p=2)``;
    ``damp(du_dt, C_max)`` returns the fixed output holding
    ``-C_max * (d / num_layers) ** p * du_dt`` on the nodes within
    ``num_layers`` of a non-excluded face and 0 elsewhere."""

    def __init__(self, inputs, num_layers: int, grid_shape, dx: float, ghost_cells: int,
                 p: int = 2, float_dtype=None) -> None:
        super().__init__(inputs, inputs, dx, ghost_cells,
                         **({} if float_dtype is None else {"float_dtype": float_dtype}))
        self.grid_shape = tuple(int(s) for s in grid_shape[:3])
        self.num_layers = num_layers
        self.p = np.float32(p)
        self.exclude_group = set()
        self.indices, self.groups = layer_groups(self.grid_shape, num_layers)

    def exclude(self, *groups: str) -> None:
        for group in groups:
            if group not in self.groups:
                raise AssertionError(f"unknown group {group!r}")
            self.exclude_group.add(group)

    def get_damping_array(self):
        return layer_points(self.indices, self.groups, self.exclude_group)

    @staticmethod
    def calculate_C_max(c: float, L: float, p: float = 2, R: float = 1e-2) -> float:
        """(p + 1) c / (2 L) ln(1 / R) (``viscousDampingLayer.py:91-105``)."""
        return float((p + 1.0) * c / (2.0 * L) * np.log(1.0 / R))

    def __call__(self, du_dt, C_max: float):
        return super().__call__(du_dt, C_max)

    @setup
    def initialize_kernel(self, du_dt, C_max: float) -> None:
        self._check_input(du_dt)
        self.outer_layers = self.get_damping_array()
        self.warp_outer_points = point_table(self.outer_layers, self.grid_shape, self.num_layers,
                                             du_dt.device)
        self.kernel = create_dampingLayer_kernel(self.input_dtype, self.num_layers, self.p,
                                                 self.grid_shape, self.ghost_cells)

    @setup(order=2)
    def zero_array(self, *args, **kwargs) -> None:
        self.output_array.zero_()

    def forward(self, du_dt, C_max: float):
        return viscousDampingLayer(self.kernel, du_dt, self.warp_outer_points, C_max,
                                   self.output_array, device=self.device)
