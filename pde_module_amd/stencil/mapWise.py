"""Map-wise module over one field (reference ``pm/stencil/mapWise.py:8-44``)."""
from __future__ import annotations

import torch

from .. import _native as N
from .element_ops import MAP_OPS, FunctionMap, MapOp
from .hooks import setup
from .stencil import Stencil


def resolve_map_op(element_op) -> MapOp:
    if isinstance(element_op, MapOp):
        return element_op
    if isinstance(element_op, str):
        if element_op in MAP_OPS:
            return MAP_OPS[element_op]
        raise TypeError(f"unknown map {element_op!r}: the compiled maps are "
                        f"{', '.join(sorted(MAP_OPS))}; any other map is given as a function")
    if callable(element_op):
        return FunctionMap(element_op)
    raise TypeError(f"a map is a compiled map's name, a MapOp or a function, not "
                    f"{type(element_op).__name__}")


class MapWise(Stencil):
    """``MapWise(element_op)``; ``module(A, *args)`` returns the fixed output
    (same shape as A) holding ``element_op(A[n], *args)``.  The extra arguments
    are the map's own (scalars, or a field for ``divide_by_scalar``).

    ``element_op`` is a compiled map (``shifted_scale``, ``divide_by_scalar``:
    one HIP kernel each) or any Python function over whole fields, element
    axes last, run as torch operations on the device (``FunctionMap``)."""

    def __init__(self, element_op) -> None:
        self.element_op = resolve_map_op(element_op)
        super().__init__()

    @setup(order=1)
    def initialize_kernel(self, array_A: torch.Tensor, *args, **kwargs) -> None:
        self.input_dtype = self.output_dtype = tuple(array_A.shape[3:])
        self.output_array = self.create_output_array(array_A)

    def forward(self, array_A: torch.Tensor, *args, **kwargs):
        N.check_field(array_A, "array_A")
        if array_A.shape != self.output_array.shape:
            raise AssertionError("array shape changed after setup")
        self.element_op.launch(array_A, self.output_array, *args)
        return self.output_array
