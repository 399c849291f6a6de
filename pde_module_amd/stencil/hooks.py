"""Hook decorators that register a Stencil method to run at setup, before or
after forward (reference ``pm/stencil/hooks.py:6-102``): lower ``order`` runs
first; the ``debug`` flag is stored but, as in the reference, not read."""
from __future__ import annotations

from typing import Callable, Optional

_TAGS = ("_setup_order", "_before_forward_order", "_after_forward_order")


def _register(tag: str, fn: Optional[Callable], order: int, debug: bool):
    if not isinstance(order, int):
        raise AssertionError("order must be integer")

    def decorator(func: Callable) -> Callable:
        if any(hasattr(func, t) for t in _TAGS):
            raise AssertionError("Method already registered!")
        setattr(func, tag, order)
        func._debug = debug
        return func

    return decorator if fn is None else decorator(fn)


def setup(fn: Optional[Callable] = None, *, order: int = 0, debug: bool = False):
    """Run the method once, on the first call of the stencil."""
    return _register("_setup_order", fn, order, debug)


def before_forward(fn: Optional[Callable] = None, *, order: int = 0, debug: bool = False):
    """Run the method before every forward."""
    return _register("_before_forward_order", fn, order, debug)


def after_forward(fn: Optional[Callable] = None, *, order: int = 0, debug: bool = False):
    """Run the method after every forward."""
    return _register("_after_forward_order", fn, order, debug)
