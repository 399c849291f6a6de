"""Device evaluation of Python boundary functions (``FunctionValue``) into the
time array the boundary kernels read: ``t[0]`` is the time, ``t[1 + slot]``
the value an array-valued op (``PDE_BCV_ARRAY``) writes.

Each function group owns a contiguous run of slots: ``M * C`` for a
functional group (node-major, the function's vector per node) and ``M`` for a
legacy group (component ``varID`` per entry).  ``FunctionSlots.evaluate``
calls every group's function once on the whole group and copies the results
into their slots, all as torch operations on the field's device and stream.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Callable

import numpy as np
import torch


@dataclass
class FunctionGroup:
    key: str                 # the params key (the group's name)
    func: Callable
    ijk: np.ndarray          # (M, 3) node indices
    base: int                # first slot
    width: int               # C (functional: the node vector) or 1 (legacy)
    var: int | None          # legacy: the component passed as varID


class FunctionSlots:
    def __init__(self, groups: list[FunctionGroup], device, dx32, coordinates=None):
        self.groups = groups
        self.nslots = sum(len(g.ijk) * g.width for g in groups)
        self.device = device
        self.dx = float(np.float32(dx32))
        self.values = torch.zeros(1 + self.nslots, dtype=torch.float32, device=device)
        self._ijk = [torch.from_numpy(np.ascontiguousarray(g.ijk, dtype=np.int64)).to(device)
                     for g in groups]
        self.coordinates = None
        if coordinates is not None:
            c = torch.as_tensor(np.asarray(coordinates) if not isinstance(coordinates, torch.Tensor)
                                else coordinates)
            self.coordinates = c.to(device=device, dtype=torch.float32)

    def evaluate(self, u: torch.Tensor, t: torch.Tensor, params: dict[str, Any] | None):
        """Writes t (a 1-element device array) and every group's values;
        returns the array to hand to the boundary kernels."""
        vals = self.values
        if t.data_ptr() != vals.data_ptr():
            vals[0:1].copy_(t)
        tt = vals[0]
        params = params or {}
        for g, ijk in zip(self.groups, self._ijk):
            M = len(g.ijk)
            if M == 0:
                continue
            p = params.get(g.key, 0)
            if g.var is None:
                r = g.func(u, ijk, self.coordinates, tt, self.dx, p)
                shape = (M, g.width)
            else:
                r = g.func(u, ijk, g.var, self.coordinates, tt, self.dx, p)
                shape = (M,)
            r = torch.as_tensor(r, dtype=torch.float32, device=self.device)
            try:
                r = r.broadcast_to(shape)
            except RuntimeError:
                raise ValueError(f"boundary function of group {g.key!r} returned shape "
                                 f"{tuple(r.shape)}, expected one broadcastable to {shape}")
            vals[1 + g.base:1 + g.base + M * g.width].view(shape).copy_(r)
        return vals
