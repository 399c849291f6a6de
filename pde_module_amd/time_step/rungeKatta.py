"""Runge-Kutta time integration of several fields (reference
``pm/time_step/rungeKatta.py:21-174``).

Both integrators take the fields' element shapes (the role of the reference's
Warp ``vec``/``mat`` dtypes: ``(C,)`` for a vec(C) field, ``(N, D)`` for a
mat(N, D) one; :func:`field_dtype` reads it off a field), a right-hand side
``f(t, *fields, **kwargs) -> tuple of rates`` and a boundary operator
``bc(t, *fields, **kwargs) -> tuple of fields``, which must have the same
signature.  Every update is a :class:`ForwardEuler` module (the HIP
``y + dt * k`` kernel); the copy of the start-of-step fields is the library's
copy kernel, so a whole step can be captured in a HIP graph.

Differences from the reference, both in code that cannot run there:

* ``check_args`` compares each field's element shape with the one given to
  the constructor (the reference reads an ``input_dtype`` that its
  ``ForwardEuler`` never sets, ``rungeKatta.py:59,132``, so its setup raises
  ``AttributeError``);
* ``RungeKatta1`` steps by ``dt``, the forward Euler its docstring describes
  (``rungeKatta.py:22-25``); the reference's code uses ``dt / 2``
  (``rungeKatta.py:80``).
"""
from __future__ import annotations

import inspect
from typing import Callable

import torch

from .. import _native as N
from ..stencil import Stencil, setup
from ..stencil.stencil import tuplify
from ..utils import SignatureMismatchError
from .forwardEuler import ForwardEuler


def field_dtype(field: torch.Tensor) -> tuple:
    """Element shape of a field: ``(C,)`` for (X, Y, Z, C), ``(N, D)`` for
    (X, Y, Z, N, D)."""
    return tuple(int(s) for s in field.shape[3:])


def _dtypes(field_dtypes) -> tuple:
    """One element shape (a tuple of ints) or a sequence of them."""
    if isinstance(field_dtypes, torch.Tensor):
        return (field_dtype(field_dtypes),)
    items = tuplify(field_dtypes)
    if items and all(isinstance(i, int) for i in items):
        return (tuple(items),)
    return tuple(field_dtype(i) if isinstance(i, torch.Tensor) else tuple(tuplify(i))
                 for i in items)


class _RungeKatta(Stencil):
    def __init__(self, field_dtypes, f: Callable, bc: Callable) -> None:
        super().__init__()
        self.input_dtypes = _dtypes(field_dtypes)
        self.f = f
        self.bc = bc
        self.num_fields = len(self.input_dtypes)
        if inspect.signature(f) != inspect.signature(bc):
            raise SignatureMismatchError(
                f"Signature For f {inspect.signature(f)} does not match the signature of "
                f"bc {inspect.signature(bc)}!")

    @setup
    def check_args(self, t: float, dt: float, *arrays: torch.Tensor, **kwargs) -> None:
        """Every field is a GPU FP32 tensor whose element shape matches the
        constructor's (``rungeKatta.py:54-60``)."""
        if len(arrays) != self.num_fields:
            raise AssertionError(f"expected {self.num_fields} fields, got {len(arrays)}")
        for arr, want in zip(arrays, self.input_dtypes):
            N.check_field(arr, "field")
            if field_dtype(arr) != want:
                raise AssertionError("arrays and input dtype for euler must match")


class RungeKatta1(_RungeKatta):
    """Forward Euler for several fields: ``bc``, then ``y + dt * f(t, y)``."""

    def __init__(self, field_dtypes, f: Callable, bc: Callable) -> None:
        super().__init__(field_dtypes, f, bc)
        self.forwardEulers = tuple(ForwardEuler() for _ in self.input_dtypes)

    def forward(self, t: float, dt: float, *arrays: torch.Tensor, **kwargs) -> tuple:
        arrays = self.bc(t, *arrays, **kwargs)
        k1s = self.f(t, *arrays, **kwargs)
        return tuple(fe(arr, k1, dt) for fe, arr, k1 in zip(self.forwardEulers, arrays, k1s))


class RungeKatta2(_RungeKatta):
    """Midpoint method for several fields (``rungeKatta.py:85-174``)::

        y0 = bc(t, y)              (copied: bc may return its own fixed outputs)
        k1 = f(t, y0)
        y1 = bc(t, y0 + dt/2 * k1)
        k2 = f(t + dt/2, y1)
        y  = y0 + dt * k2

    Returns the second set of Euler modules' fixed outputs."""

    def __init__(self, field_dtypes, f: Callable, bc: Callable) -> None:
        super().__init__(field_dtypes, f, bc)
        self.forwardEulers_1 = tuple(ForwardEuler() for _ in self.input_dtypes)
        self.forwardEulers_2 = tuple(ForwardEuler() for _ in self.input_dtypes)

    @setup
    def set_buffers(self, t: float, dt: float, *arrays: torch.Tensor, **kwargs) -> None:
        """Buffers for the start-of-step fields (``rungeKatta.py:140-143``)."""
        self.t0_fields = tuple(torch.empty_like(arr) for arr in arrays)

    def _copy_to_buffers(self, arrays) -> None:
        for buf, arr in zip(self.t0_fields, arrays):
            if arr.shape != buf.shape:
                raise AssertionError("field shape changed after setup")
            N.call("pde_copy_f32", N.dev_ptr(arr), N.dev_ptr(buf), arr.numel(), N.stream_of(arr))

    def forward(self, t: float, dt: float, *arrays: torch.Tensor, **kwargs) -> tuple:
        arrays = self.bc(t, *arrays, **kwargs)
        self._copy_to_buffers(arrays)
        k1s = self.f(t, *arrays, **kwargs)
        y1s = [fe(arr, k1, dt / 2) for fe, arr, k1 in zip(self.forwardEulers_1, arrays, k1s)]
        y1s = self.bc(t, *y1s, **kwargs)
        k2s = self.f(t + dt / 2, *y1s, **kwargs)
        return tuple(fe(arr, k2, dt)
                     for fe, arr, k2 in zip(self.forwardEulers_2, self.t0_fields, k2s))
