"""Element-wise FDM modules (reference ``pm/FDM/module/elementWiseOps.py:10-89``)."""
from __future__ import annotations

import torch

from ...stencil.elementWise import ElementWise
from ...stencil.element_ops import outer, scalar_vector_multiply
from ...stencil.hooks import setup
from ...stencil.stencil import tuplify
from .. import functional


class OuterProduct(ElementWise):
    """``OuterProduct(vec_A, vec_B)``: two vector fields -> mat(vec_A, vec_B)
    field, M[i][j] = A[i] * B[j].

    The reference's dtype check compares B against vec_A
    (``pm/FDM/module/elementWiseOps.py:27-29``), so it only accepts
    vec_A == vec_B; here B is checked against vec_B, which agrees with it on
    every input it accepts."""

    def __init__(self, vec_A: int, vec_B: int, float_dtype=torch.float32) -> None:
        super().__init__(outer, output_dtype=(int(vec_A), int(vec_B)))
        self.vec_A, self.vec_B = (int(vec_A),), (int(vec_B),)

    @setup(order=2)
    def check_same(self, array_A, array_B, *args, **kwargs) -> None:
        if tuple(array_A.shape[3:]) != self.vec_A or tuple(array_B.shape[3:]) != self.vec_B:
            raise AssertionError(f"expected vec{self.vec_A[0]} and vec{self.vec_B[0]} fields")

    def forward(self, array_A, array_B, *args, **kwargs):
        return functional.elementWiseOp(self.element_op, array_A, array_B, self.output_array)


class scalarVectorMult(ElementWise):
    """``scalarVectorMult(outputs)``: a vec(1) field times a vector / matrix
    field whose element shape is ``outputs`` (int or tuple)."""

    def __init__(self, outputs, float_dtype=torch.float32) -> None:
        super().__init__(scalar_vector_multiply, output_dtype=tuplify(outputs))

    @setup(order=2)
    def check_same(self, array_A, array_B, *args, **kwargs) -> None:
        if tuple(array_A.shape[3:]) not in ((1,), ()) or \
                tuple(array_B.shape[3:]) != self.output_dtype:
            raise AssertionError("scalarVectorMult needs a vec(1) field and a "
                                 f"{self.output_dtype} field")

    def forward(self, array_A, array_B, *args, **kwargs):
        return functional.elementWiseOp(self.element_op, array_A, array_B, self.output_array)
