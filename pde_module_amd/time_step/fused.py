"""Fused explicit steps: boundary + Laplacian + update in one pass over HBM.

The reference runs these loops as separate modules (``examples/2D_Heat_Equation.py:58-62``:
BC copy, Laplacian, ForwardEuler; ``examples/wave2d/standingWave.py:72-76``: the
same plus a second Euler), moving about 28 bytes per lattice update for a scalar
heat step.  Here the boundary writes go into the state in place (a surface-sized
launch) and one kernel reads the state once and writes the next one: 8 bytes per
update for heat, 20 for the u/v wave step.  On 2D scalar fields the heat step
folds the boundary into the same launch (``pde_heat_step_ring_f32``: the outer
rows and columns apply it in LDS) whenever the boundary table allows, so a
step is a single launch.  Every non-ghost value is the same FP32 number the
module-by-module loop gives.

The steppers own two ping-pong buffers.  The first call copies the caller's
state into them; later calls are expected to pass back the tensor the previous
call returned (``u = step(u)``), exactly like the reference loop's
``u = u_next``.  The state passed in is modified (its boundary is applied).
"""
from __future__ import annotations

import numpy as np
import torch

from .. import _native as N
from ..FDM.module.laplacian import laplacian_weights


class HeatStep:
    """``u_next = u2 + dt * (alpha * lap(u2))`` with ``u2 = boundary(u)``.

    ``boundary`` is a GridBoundary (functional or legacy) built for the field."""

    def __init__(self, boundary, dx: float, ghost_cells: int, alpha: float, dt: float,
                 one_launch: bool = True) -> None:
        self.boundary = boundary
        self.weights = laplacian_weights(dx)
        self.ghost_cells = ghost_cells
        self.alpha = np.float32(alpha)
        self.dt = np.float32(dt)
        self._bufs = None
        # one_launch=False keeps the boundary launch and the step launch apart
        # (same values; used to check the one-launch form)
        self.one_launch = bool(one_launch)
        # (key, DeviceRingPlan or None, table): the table is held so that its id,
        # part of the key, cannot be reused by a later table while cached
        self._ring = None
        # the same for the multi-step tile plan: (key, device arrays or None, table)
        self._mplan = None

    def _multistep_plan(self, u: torch.Tensor, halo: int, region=None):
        """Device tiles, tile ops and grid record of the multi-step launch for
        u's shape, the boundary's current table and this halo, or None when
        the field or its boundary cannot take it (2D scalar fields only)."""
        shape = tuple(int(s_) for s_ in u.shape[:3])
        if u.shape[3] != 1 or sum(1 for s_ in shape if s_ > 1) != 2:
            return None
        self.boundary._ensure_table(u)
        table = self.boundary.table
        p1, p2 = [s_ for s_ in shape if s_ > 1]
        if region is None:
            # a small grid steps fastest as many small tiles (one workgroup
            # each; measured on 258^2: 48 x 48 regions at 12 steps per launch
            # 1.43 us/step, 128 x 128 2.7); large grids take full-size regions
            region = (128, 128)
            if p1 * p2 <= (1 << 18):
                region = next(((c, c) for c in (48, 64, 96) if c > 2 * halo + 16), region)
        key = (shape, u.device, id(table), int(halo), tuple(region))
        if self._mplan is None or self._mplan[0] != key:
            from ..multistep import PlanError, check_coverage, plan_grid
            try:
                tiles, tops = plan_grid(p1, p2, table.host_ops, halo, rows_cap=region[0],
                                        cols_cap=region[1])
                check_coverage(tiles, [p1], [p2])
            except PlanError:
                self._mplan = (key, None, table)
                return None
            grid = np.zeros(1, N.BATCH_GRID_DTYPE)
            wm, wc, wp = self.weights
            grid[0] = (0, 0, 0, p1, p2, self.ghost_cells, 0, wm, wc, wp, self.alpha, self.dt, 0.0)
            dev = lambda a: torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).copy()).to(
                u.device)
            if not len(tops):
                tops = np.zeros(1, N.HEAT_TILE_OP_DTYPE)
            self._mplan = (key, (dev(tiles), len(tiles), dev(tops), dev(grid)), table)
        return self._mplan[1]

    def run(self, u: torch.Tensor, steps: int, t=None, steps_per_launch: int = 12,
            region=None) -> torch.Tensor:
        """``steps`` steps from u with the same time t for every step: the
        values of calling the step ``steps`` times.  On 2D scalar fields whose
        boundary allows it, ``steps_per_launch`` steps run per launch on
        LDS-resident tiles (``pde_heat_multistep_f32``: each launch reads the
        state once and writes it once); otherwise one step per launch.
        ``region``: (rows, columns) cap of the tiles' LDS regions (at most
        128 x 128; default: for grids up to 2^18 cells the smallest of 48, 64,
        96 that exceeds 2 * steps_per_launch + 16, else 128 x 128).  Returns the final state (one of the stepper's buffers); u may
        be modified."""
        N.check_field(u, "u")
        steps = int(steps)
        u = self._own(u)
        spl = int(steps_per_launch)
        plan = self._multistep_plan(u, spl, region) if spl > 1 and steps > 1 else None
        if plan is None:
            for _ in range(steps):
                u = self(u, t)
            return u
        tiles, nt, tops, grid = plan
        tt = self.boundary._time_tensor(u, t)
        table = self.boundary.table
        while steps > 0:
            n = min(spl, steps)
            nxt = self._bufs[1] if u.data_ptr() == self._bufs[0].data_ptr() else self._bufs[0]
            N.call("pde_heat_multistep_f32", N.dev_ptr(u), N.dev_ptr(nxt), grid.data_ptr(), 1,
                   tiles.data_ptr(), nt, tops.data_ptr(),
                   table.ops.data_ptr(), tt.data_ptr() if tt is not None else None, n,
                   N.stream_of(u))
            u = nxt
            steps -= n
        return u

    def _ring_plan(self, u: torch.Tensor):
        """The ring plan of the one-launch step for u's shape and the
        boundary's current table, or None when the step cannot take it."""
        shape = tuple(int(s_) for s_ in u.shape[:3])
        if not self.one_launch or u.shape[3] != 1 or sum(1 for s_ in shape if s_ > 1) != 2:
            return None
        self.boundary._ensure_table(u)
        table = self.boundary.table
        key = (shape, u.device, id(table))
        if self._ring is None or self._ring[0] != key:
            from ..ring_plan import DeviceRingPlan, build_ring_plan, check_step_sources
            p1, p2 = [s_ for s_ in shape if s_ > 1]
            try:
                plan = build_ring_plan(table.host_ops, p1, p2, self.ghost_cells)
                check_step_sources(plan, table.host_ops, p1, p2)
                dev_plan = DeviceRingPlan(plan, u.device)
            except ValueError:
                dev_plan = None
            self._ring = (key, dev_plan, table)
        return self._ring[1]

    def _own(self, u: torch.Tensor) -> torch.Tensor:
        if self._bufs is None or self._bufs[0].shape != u.shape or self._bufs[0].device != u.device:
            self._bufs = (torch.empty_like(u), torch.empty_like(u))
        a, b = self._bufs
        if u.data_ptr() in (a.data_ptr(), b.data_ptr()):
            return u
        a.copy_(u)
        return a

    def step_into(self, u: torch.Tensor, out: torch.Tensor, t=None) -> None:
        """One step from u into out (distinct tensors of u's shape); u is left
        holding the boundary-applied state u2."""
        X, Y, Z = (int(s) for s in u.shape[:3])
        wm, wc, wp = self.weights
        ring = self._ring_plan(u)
        if ring is not None:
            table = self.boundary.table
            tt = self.boundary._time_tensor(u, t)
            N.call("pde_heat_step_ring_f32", N.dev_ptr(u), N.dev_ptr(out), X, Y, Z,
                   self.ghost_cells, table.ops.data_ptr() if table.nops else None, table.nops,
                   tt.data_ptr() if tt is not None else None, wm, wc, wp, self.alpha, self.dt,
                   ring.ptr, N.stream_of(u))
            return
        self.boundary.apply_inplace(u, t)
        N.call("pde_heat_step_f32", N.dev_ptr(u), N.dev_ptr(out), X, Y, Z, int(u.shape[3]),
               self.ghost_cells, wm, wc, wp, self.alpha, self.dt, N.stream_of(u))

    def __call__(self, u: torch.Tensor, t=None) -> torch.Tensor:
        N.check_field(u, "u")
        u = self._own(u)
        nxt = self._bufs[1] if u.data_ptr() == self._bufs[0].data_ptr() else self._bufs[0]
        self.step_into(u, nxt, t)
        return nxt


class WaveStep:
    """u/v wave step with the reference loop's aliasing:
    ``v_next = v + dt * (c2 * lap(u2))``, ``u_next = u2 + dt * v'`` where v' is
    the old v on the first call and v_next afterwards (symplectic Euler).
    ``c2`` is a number (used as the Laplacian's alpha) or a per-node (X, Y, Z, 1)
    tensor multiplying ``lap(u2)`` computed with alpha = 1."""

    def __init__(self, boundary, dx: float, ghost_cells: int, c2, dt: float) -> None:
        self.boundary = boundary
        self.weights = laplacian_weights(dx)
        self.ghost_cells = ghost_cells
        self.dt = np.float32(dt)
        if isinstance(c2, torch.Tensor):
            N.check_field(c2, "c2")
            self.c2, self.alpha = c2, np.float32(1.0)
        else:
            self.c2, self.alpha = None, np.float32(c2)
        self.first = True
        self._u = None
        self._v = None

    def __call__(self, u: torch.Tensor, v: torch.Tensor, t=None):
        N.check_field(u, "u")
        N.check_field(v, "v")
        if self._u is None or self._u[0].shape != u.shape:
            self._u = (torch.empty_like(u), torch.empty_like(u))
            self._v = torch.empty_like(v)
        a, b = self._u
        if u.data_ptr() not in (a.data_ptr(), b.data_ptr()):
            a.copy_(u)
            u = a
        if v.data_ptr() != self._v.data_ptr():
            self._v.copy_(v)
        nxt = b if u.data_ptr() == a.data_ptr() else a
        self.boundary.apply_inplace(u, t)
        X, Y, Z = (int(s) for s in u.shape[:3])
        wm, wc, wp = self.weights
        c2p = N.dev_ptr(self.c2) if self.c2 is not None else None
        N.call("pde_wave_step_f32", N.dev_ptr(u), N.dev_ptr(self._v), c2p, N.dev_ptr(nxt),
               N.dev_ptr(self._v), X, Y, Z, self.ghost_cells, wm, wc, wp, self.alpha, self.dt,
               1 if self.first else 0, N.stream_of(u))
        self.first = False
        return nxt, self._v
