"""Element-wise module between two fields (reference
``pm/stencil/elementWise.py:7-81``)."""
from __future__ import annotations

from typing import Optional

import torch

from .. import _native as N
from .element_ops import ELEMENT_OPS, ElementOp, FunctionElementOp, _elem
from .hooks import setup
from .stencil import Stencil


def resolve_element_op(element_op) -> ElementOp:
    if isinstance(element_op, ElementOp):
        return element_op
    if isinstance(element_op, str):
        if element_op in ELEMENT_OPS:
            return ELEMENT_OPS[element_op]
        raise TypeError(f"unknown element op {element_op!r}: the compiled ones are "
                        f"{', '.join(sorted(ELEMENT_OPS))}; any other is given as a function")
    if callable(element_op):
        return FunctionElementOp(element_op)
    raise TypeError(f"an element op is a compiled op's name, an ElementOp or a function, not "
                    f"{type(element_op).__name__}")


class ElementWise(Stencil):
    """``ElementWise(element_op, output_dtype=None)``; ``module(A, B)`` returns
    the fixed output holding ``element_op(A[n], B[n])`` at every padded node.

    ``element_op`` is one of ``pde_module_amd.stencil.element_ops``' compiled
    ops (or its name), or any Python function over whole fields, element axes
    last, run as torch operations on the device; ``output_dtype`` is the
    output's element shape (default: the one the op produces; A's for a
    function, as the reference defaults to A's dtype)."""

    def __init__(self, element_op, output_dtype: Optional[tuple] = None) -> None:
        self.element_op = resolve_element_op(element_op)
        self.output_dtype = tuple(output_dtype) if output_dtype is not None else None
        super().__init__()

    @setup(order=1)
    def initialize_kernel(self, array_A: torch.Tensor, array_B: torch.Tensor, *args,
                          **kwargs) -> None:
        if tuple(array_A.shape[:3]) != tuple(array_B.shape[:3]):
            raise AssertionError("both fields must have the same grid shape")
        self.array_A_dtype, self.array_B_dtype = _elem(array_A), _elem(array_B)
        produced = self.element_op.out_shape(self.array_A_dtype, self.array_B_dtype)
        if self.output_dtype is None:
            self.output_dtype = produced
        elif isinstance(self.element_op, FunctionElementOp):
            pass                                  # the function decides its shape
        elif self.output_dtype != produced:
            raise AssertionError(f"output dtype {self.output_dtype} but {self.element_op.name} "
                                 f"produces {produced}")
        self.output_array = self.create_output_array(array_A, self.output_dtype)

    def forward(self, array_A: torch.Tensor, array_B: torch.Tensor, *args, **kwargs):
        N.check_field(array_A, "array_A")
        N.check_field(array_B, "array_B")
        self.element_op.launch(array_A, array_B, self.output_array)
        return self.output_array
