"""Module lifecycle shared by every operator (reference
``pm/stencil/stencil.py:23-132``).

A Stencil behaves like a ``torch.nn.Module`` with a fixed output: the first call
runs the setup hooks (which allocate the output once), every call then runs the
before-forward hooks, ``forward`` and the after-forward hooks, and returns the
same output tensor each time, overwritten in place.  Outputs come from
``torch.empty`` and are therefore not zeroed, like the reference's ``wp.empty``.
"""
from __future__ import annotations

from collections.abc import Iterable
from typing import Any, Optional

import torch


def tuplify(x: Any) -> tuple:
    """Single item -> 1-tuple; strings/bytes count as one item (reference
    ``pm/utils/utils.py:21-32``)."""
    if isinstance(x, (str, bytearray, bytes)):
        return (x,)
    if isinstance(x, Iterable):
        return tuple(x)
    return (x,)


class Stencil:
    """Array-of-structures stencil base module.  Subclass and implement
    ``forward``; decorate set-up and per-call work with the hooks in
    ``pde_module_amd.stencil.hooks``."""

    initial: bool = True
    device: Optional[torch.device] = None

    def __init__(self, *args, **kwargs) -> None:
        pass

    @staticmethod
    def create_output_array(input_array: torch.Tensor,
                            element_shape: Optional[tuple] = None) -> torch.Tensor:
        """Unzeroed output on the input's device: the input's grid shape
        (first three axes) followed by ``element_shape`` (default: the input's
        own element shape)."""
        grid = tuple(input_array.shape[:3])
        elem = tuple(input_array.shape[3:]) if element_shape is None else tuple(element_shape)
        return torch.empty(grid + elem, dtype=input_array.dtype, device=input_array.device)

    def forward(self, *args, **kwargs) -> Any:
        raise NotImplementedError(
            "a Stencil subclass defines forward()")

    def _set_registry(self, call: str) -> str:
        name = f"_{call}_list"
        if not hasattr(self, name):
            found = []
            for attr in dir(self):
                method = getattr(self, attr)
                if callable(method) and hasattr(method, f"_{call}_order"):
                    found.append(method)
            setattr(self, name, sorted(found, key=lambda m: getattr(m, f"_{call}_order")))
        return name

    def setup(self, *args, **kwargs) -> None:
        """Run the setup hooks now (the first ``__call__`` does it otherwise)."""
        self._set_registry("before_forward")
        self._set_registry("after_forward")
        for method in getattr(self, self._set_registry("setup")):
            method(*args, **kwargs)
        self.initial = False

    def before_forward(self, *args, **kwargs) -> None:
        for method in getattr(self, self._set_registry("before_forward")):
            method(*args, **kwargs)

    def after_forward(self, *args, **kwargs) -> None:
        for method in getattr(self, self._set_registry("after_forward")):
            method(*args, **kwargs)

    def __call__(self, *args, **kwargs) -> Any:
        if self.initial:
            self.setup(*args, **kwargs)
        self.before_forward(*args, **kwargs)
        output = self.forward(*args, **kwargs)
        self.after_forward(*args, **kwargs)
        return output
