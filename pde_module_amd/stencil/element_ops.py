"""Node-wise element operations compiled into libpdehip.so.

The reference builds element-wise kernels from arbitrary Warp functions
(``pm/stencil/elementWise.py:57-81``, ``pm/stencil/mapWise.py:100-107``).  There
is no Warp here and no run-time code generation: the operations its modules and
examples use are a closed set of HIP kernels, named below.  Passing anything
else to ``ElementWise`` / ``MapWise`` raises ``TypeError``.

Element shapes are the trailing dims of a field: a vec(C) field is
``(X, Y, Z, C)`` -> element ``(C,)``; a mat(N, D) field ``(X, Y, Z, N, D)`` ->
``(N, D)``.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .. import _native as N


def _nodes(x: torch.Tensor) -> int:
    return int(math.prod(x.shape[:3]))


def _elem(x: torch.Tensor) -> tuple:
    return tuple(int(s) for s in x.shape[3:])


def _len(shape: tuple) -> int:
    return int(math.prod(shape)) if shape else 1


class ElementOp:
    """A binary node-wise operation ``out[n] = op(A[n], B[n])``."""

    name = "element op"

    def out_shape(self, ea: tuple, eb: tuple) -> tuple:
        raise NotImplementedError

    def launch(self, a: torch.Tensor, b: torch.Tensor, out: torch.Tensor) -> None:
        raise NotImplementedError

    def __repr__(self) -> str:
        return f"<element op {self.name}>"


class _Outer(ElementOp):
    """``wp.outer(a, b)``: mat(la, lb) with M[i][j] = a[i] * b[j]."""

    name = "outer"

    def out_shape(self, ea, eb):
        if len(ea) != 1 or len(eb) != 1:
            raise AssertionError("outer product takes two vector fields")
        return (ea[0], eb[0])

    def launch(self, a, b, out):
        N.call("pde_outer_f32", N.dev_ptr(a), N.dev_ptr(b), N.dev_ptr(out), _nodes(a),
               _elem(a)[0], _elem(b)[0], N.stream_of(a))


class _ScalarVectorMultiply(ElementOp):
    """``a[0] * b`` for a vec(1) field a and any vector/matrix field b
    (reference ``pm/FDM/kernel/elementWiseOps.py:6-23``)."""

    name = "scalar_vector_multiply"

    def out_shape(self, ea, eb):
        if _len(ea) != 1:
            raise AssertionError("the scalar operand must be a vec(1) field")
        return eb

    def launch(self, a, b, out):
        N.call("pde_scalar_mul_f32", N.dev_ptr(a), N.dev_ptr(b), N.dev_ptr(out), _nodes(a),
               _len(_elem(b)), N.stream_of(a))


class _MatVec(ElementOp):
    """Warp's ``mat * vec`` (the array product ``du * u``): out[i] =
    sum_j m[i][j] * v[j], j ascending."""

    name = "matvec"

    def out_shape(self, ea, eb):
        if len(ea) != 2 or len(eb) != 1 or ea[1] != eb[0]:
            raise AssertionError(f"matvec needs mat(N, D) and vec(D); got {ea} and {eb}")
        return (ea[0],)

    def launch(self, a, b, out):
        rows, cols = _elem(a)
        N.call("pde_matvec_f32", N.dev_ptr(a), N.dev_ptr(b), N.dev_ptr(out), _nodes(a), rows,
               cols, N.stream_of(a))


class _LinComb2(ElementOp):
    """``a + s * b`` with s = +1 (add) or -1 (sub): Warp's array ``+`` / ``-``."""

    def __init__(self, name, sign):
        self.name, self.sign = name, float(sign)

    def out_shape(self, ea, eb):
        if ea != eb:
            raise AssertionError(f"{self.name} needs equal element shapes; got {ea} and {eb}")
        return ea

    def launch(self, a, b, out):
        N.call("pde_lincomb_f32", N.dev_ptr(a), N.dev_ptr(b), None, None, np.float32(1.0),
               np.float32(self.sign), np.float32(0.0), np.float32(0.0), 2, N.dev_ptr(out),
               a.numel(), N.stream_of(a))


outer = _Outer()
scalar_vector_multiply = _ScalarVectorMultiply()
matvec = _MatVec()
add = _LinComb2("add", 1.0)
sub = _LinComb2("sub", -1.0)

ELEMENT_OPS = {op.name: op for op in (outer, scalar_vector_multiply, matvec, add, sub)}


class MapOp:
    """A node-wise map ``out[n] = f(A[n], *args)``; ``args`` are given per call."""

    name = "map"

    def launch(self, a: torch.Tensor, out: torch.Tensor, *args) -> None:
        raise NotImplementedError

    def __repr__(self) -> str:
        return f"<map {self.name}>"


class _ShiftedScale(MapOp):
    """``f(x, scale, shift) = scale * (x - shift)`` on every component, or on
    component ``comp`` only (the others copied).  The reference's
    ``get_p_op`` (rho -> c2 * (rho - density), ``examples/LDC/LDC_RK2.py:111-116``)
    is ``shifted_scale`` on a vec(1) field."""

    name = "shifted_scale"

    def __init__(self, comp: int = -1):
        self.comp = int(comp)

    def launch(self, a, out, scale, shift=0.0):
        C = _len(_elem(a))
        N.call("pde_affine_f32", N.dev_ptr(a), N.dev_ptr(out), _nodes(a), C, self.comp,
               np.float32(scale), np.float32(shift), N.stream_of(a))


class _DivideByScalar(MapOp):
    """``f(m, rho) = m / rho[0]`` for a vec(1) field rho (the reference's
    ``MapWise(lambda m, rho: m / rho[0])``, ``examples/LDC/LDC_RK2.py:117``)."""

    name = "divide_by_scalar"

    def launch(self, a, out, rho):
        N.check_field(rho, "rho")
        if _len(_elem(rho)) != 1 or _nodes(rho) != _nodes(a):
            raise AssertionError("divisor must be a vec(1) field on the same grid")
        N.call("pde_div_scalar_f32", N.dev_ptr(a), N.dev_ptr(rho), N.dev_ptr(out), _nodes(a),
               _len(_elem(a)), N.stream_of(a))


shifted_scale = _ShiftedScale()
divide_by_scalar = _DivideByScalar()


def component_shifted_scale(comp: int) -> MapOp:
    """``shifted_scale`` applied to one component."""
    return _ShiftedScale(comp)


MAP_OPS = {op.name: op for op in (shifted_scale, divide_by_scalar)}


def _copy_result(r, out: torch.Tensor, name: str) -> None:
    r = torch.as_tensor(r, dtype=out.dtype, device=out.device)
    try:
        out.copy_(r.broadcast_to(out.shape))
    except RuntimeError:
        raise AssertionError(f"{name} returned shape {tuple(r.shape)}; the output is "
                             f"{tuple(out.shape)}")


class FunctionMap(MapOp):
    """A user map written over whole fields: ``func(A, *args)`` with the
    element axes last (so ``rho[..., 0:1]`` is the reference's ``rho[0]``),
    evaluated as torch operations on the field's device and stream into the
    module's fixed output.  The reference compiles the element function into a
    Warp kernel (``pm/stencil/mapWise.py:28-30``)."""

    def __init__(self, func):
        self.func = func
        self.name = getattr(func, "__name__", "function")

    def launch(self, a, out, *args):
        _copy_result(self.func(a, *args), out, f"map {self.name}")


class FunctionElementOp(ElementOp):
    """A user element operation written over whole fields: ``func(A, B)``
    with the element axes last, evaluated as torch operations on the device
    into the module's fixed output (reference ``pm/stencil/elementWise.py:61-81``
    builds a Warp kernel calling it per node).  The output's element shape is
    A's unless the module is given ``output_dtype``."""

    def __init__(self, func):
        self.func = func
        self.name = getattr(func, "__name__", "function")

    def out_shape(self, ea, eb):
        return ea

    def launch(self, a, b, out):
        _copy_result(self.func(a, b), out, f"element op {self.name}")


def linear_combination(out: torch.Tensor, *terms) -> torch.Tensor:
    """``out = c0*x0 + c1*x1 + ...`` for up to four ``(c, x)`` terms in one
    pass, accumulated left to right; with coefficients +-1 it equals the Warp
    array-operator chain ``x0 - x1 + x2 ...`` bit for bit.  ``out`` may be one
    of the x."""
    if not 1 <= len(terms) <= 4:
        raise ValueError("linear_combination takes 1 to 4 (coefficient, field) terms")
    xs, cs = [], []
    for c, x in terms:
        N.check_field(x, "x")
        if x.shape != out.shape:
            raise AssertionError(f"shape mismatch: {tuple(x.shape)} vs {tuple(out.shape)}")
        xs.append(N.dev_ptr(x))
        cs.append(np.float32(c))
    N.check_field(out, "out")
    k = len(terms)
    xs += [None] * (4 - k)
    cs += [np.float32(0.0)] * (4 - k)
    N.call("pde_lincomb_f32", *xs, *cs, k, N.dev_ptr(out), out.numel(), N.stream_of(out))
    return out
