"""Sponge layer (reference ``pm/FDM/module/farfield.py:13-166``)."""
from __future__ import annotations

from ...stencil import setup
from ..functional import farfield
from ..kernel import create_spongeLayer_kernel
from ..layers import layer_groups, layer_points, point_table
from .explicitUniformGridStencil import ExplicitUniformGridStencil


class FarField(ExplicitUniformGridStencil):
    """``FarField(inputs, num_layers, beta, grid_shape, dx, ghost_cells)``;
    ``ff(u, farfield_condition, sigma_max)`` returns the fixed output holding
    ``-sigma_max * (d / num_layers) ** beta * (u - farfield_condition)`` on the
    nodes within ``num_layers`` of a non-excluded face (d = num_layers minus the
    distance to the nearest face) and 0 elsewhere (zeroed once at setup).
    Groups ``-X, +X, -Y, ...`` can be left out with ``exclude``."""

    def __init__(self, inputs, num_layers: int, beta: float, grid_shape, dx: float,
                 ghost_cells: int, float_dtype=None) -> None:
        super().__init__(inputs, inputs, dx, ghost_cells,
                         **({} if float_dtype is None else {"float_dtype": float_dtype}))
        self.grid_shape = tuple(int(s) for s in grid_shape[:3])
        self.num_layers = num_layers
        self.beta = beta
        self.exclude_group = set()
        self.indices, self.groups = layer_groups(self.grid_shape, num_layers)

    def exclude(self, *groups: str) -> None:
        """Leave groups (e.g. an inlet face) out of the sponge zone."""
        for group in groups:
            if group not in self.groups:
                raise AssertionError(f"unknown group {group!r}")
            self.exclude_group.add(group)

    def get_sponge_array(self):
        """Grid indices of the sponge nodes (``farfield.py:76-89``)."""
        return layer_points(self.indices, self.groups, self.exclude_group)

    def __call__(self, input_array, farfield_condition, sigma_max: float):
        return super().__call__(input_array, farfield_condition, sigma_max)

    @setup
    def initialize_kernel(self, input_array, *args, **kwargs) -> None:
        self._check_input(input_array)
        self.outer_layers = self.get_sponge_array()
        self.warp_outer_points = point_table(self.outer_layers, self.grid_shape, self.num_layers,
                                             input_array.device)
        self.kernel = create_spongeLayer_kernel(self.input_dtype, self.num_layers, self.beta,
                                                self.grid_shape, self.ghost_cells)

    @setup(order=2)
    def zero_array(self, *args, **kwargs) -> None:
        self.output_array.zero_()

    def forward(self, input_array, farfield_condition, sigma_max: float):
        return farfield(self.kernel, input_array, self.warp_outer_points, farfield_condition,
                        sigma_max, self.output_array, device=self.device)
