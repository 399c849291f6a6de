"""Incompressible lid-driven cavity by pressure projection (2D, collocated).

The reference's cavity examples use the artificial-compressibility method
(``examples/LDC/LDC_No_Graph.py``, ``LDC_RK2.py``); this is the projection
variant of the same flow, built from the same operators:

    u2  = B_u(u)                                  no slip, lid u = U on +Y
    u*  = u2 + dt * (-(grad u2) u2 + nu * lap(u2))
    u*  = B_u(u*)
    rhs = div(u*) / dt
    p   <- J(B_p(p), rhs), sweeps times           Jacobi sweeps, Neumann-0 B_p
    u   = u* - dt * grad(B_p(p))

with J one Jacobi sweep of lap(p) = rhs (``pde_jacobi_f32``).  ``step`` runs on
preallocated buffers through the HIP kernels (capturable in a HIP graph).
``step_adjoint`` is its transpose: the same kernels' transposes in reverse
order on preallocated buffers (also capturable); the step is nonlinear only
through u (the convection term), so the transpose needs just the step's input
velocity.  ``kinetic_energy_rollout`` gives dJ/du0 (and dJ/dp0) for J = 1/2
sum |u|^2 dx^2 over the interior at the final step through ``CavityRollout``
(every step's velocity kept, or checkpoints every ``checkpoint_every`` steps
and a recomputed segment); ``kinetic_energy_rollout_autograd`` is the same
gradient through torch autograd over the differentiable operators
(``diff_step``), the cross-check.
"""
from __future__ import annotations

import numpy as np
import torch
import torch.utils.checkpoint as cp

from . import _native as N
from . import autograd as A
from .deprecated import GridBoundary as LegacyGridBoundary
from .FDM.boundary.device_table import DeviceBoundaryTable
from .FDM.module.grad import central_weights
from .FDM.module.laplacian import laplacian_weights

# the most Jacobi sweeps one multi-sweep launch runs
MAX_SWEEPS_PER_LAUNCH = N.JACOBI_MAX_SWEEPS


def _grid(x):
    return tuple(int(s) for s in x.shape[:3])


def _repeat_table(table, batch, stride, device):
    """A boundary table for ``batch`` instances of one grid stored one after
    another: ``table``'s ops repeated per instance, their scalar indices
    shifted by the instance's offset (``stride`` scalars per instance), in
    the same order within each instance."""
    ops = table.host_ops
    reps = []
    for i in range(batch):
        o = ops.copy()
        o["dst"] += i * stride
        rd = o["kind"] != N.BC_SET
        o["src"][rd] += i * stride
        reps.append(o)
    return DeviceBoundaryTable(np.concatenate(reps) if reps else ops, device)


class _MatVec(torch.autograd.Function):
    """out[n] = M[n] v[n]; backward gM = outer(g, v), gv = M^T g (HIP kernels)."""

    @staticmethod
    def forward(ctx, m, v):
        out = torch.empty(tuple(m.shape[:4]), dtype=m.dtype, device=m.device)
        nodes = int(np.prod(m.shape[:3]))
        N.call("pde_matvec_f32", m.data_ptr(), v.data_ptr(), out.data_ptr(), nodes,
               int(m.shape[3]), int(m.shape[4]), N.stream_of(m))
        ctx.save_for_backward(m, v)
        return out

    @staticmethod
    def backward(ctx, g):
        m, v = ctx.saved_tensors
        g = g.contiguous()
        nodes = int(np.prod(m.shape[:3]))
        rows, cols = int(m.shape[3]), int(m.shape[4])
        gm = torch.empty_like(m)
        N.call("pde_outer_f32", g.data_ptr(), v.data_ptr(), gm.data_ptr(), nodes, rows, cols,
               N.stream_of(g))
        gv = torch.empty_like(v)
        N.call("pde_matvec_t_f32", m.data_ptr(), g.data_ptr(), gv.data_ptr(), nodes, rows, cols,
               N.stream_of(g))
        return gm, gv


class _LinComb(torch.autograd.Function):
    """c0 x0 + c1 x1 (+ ...) left to right (pde_lincomb_f32); backward scales."""

    @staticmethod
    def forward(ctx, coeffs, *xs):
        out = torch.empty_like(xs[0])
        ptrs = [x.data_ptr() for x in xs] + [None] * (4 - len(xs))
        cs = [np.float32(c) for c in coeffs] + [np.float32(0)] * (4 - len(coeffs))
        N.call("pde_lincomb_f32", *ptrs, *cs, len(xs), out.data_ptr(), out.numel(),
               N.stream_of(out))
        ctx.coeffs = coeffs
        return out

    @staticmethod
    def backward(ctx, g):
        g = g.contiguous()
        grads = []
        for c in ctx.coeffs:
            gx = torch.empty_like(g)
            N.call("pde_scale_f32", g.data_ptr(), gx.data_ptr(), g.numel(), np.float32(c),
                   N.stream_of(g))
            grads.append(gx)
        return (None, *grads)


class _Jacobi(torch.autograd.Function):
    @staticmethod
    def forward(ctx, p, rhs, g, h2, inv):
        out = torch.zeros_like(p)
        N.call("pde_jacobi_f32", p.data_ptr(), rhs.data_ptr(), out.data_ptr(), *_grid(p), g,
               np.float32(h2), np.float32(inv), N.stream_of(p))
        ctx.meta = (g, h2, inv)
        return out

    @staticmethod
    def backward(ctx, gout):
        g, h2, inv = ctx.meta
        gout = gout.contiguous()
        gp = torch.empty_like(gout)
        grhs = torch.empty_like(gout)
        N.call("pde_jacobi_adjoint_f32", gout.data_ptr(), gp.data_ptr(), grhs.data_ptr(),
               *_grid(gout), g, np.float32(h2), np.float32(inv), N.stream_of(gout))
        return gp, grhs, None, None, None


class _KineticEnergy(torch.autograd.Function):
    """1/2 sum over interior nodes of |u|^2 dx^2 (torch reduction of the
    interior squares; the gradient is dx^2 u on the interior, 0 on ghosts)."""

    @staticmethod
    def forward(ctx, u, mask, area):
        ctx.save_for_backward(u, mask)
        ctx.area = area
        return 0.5 * area * (u * u * mask).sum()

    @staticmethod
    def backward(ctx, g):
        u, mask = ctx.saved_tensors
        return g * ctx.area * u * mask, None, None


class ProjectionCavity:
    """``ProjectionCavity(n, nu=0.01, U=1.0, dt=None, jacobi_sweeps=50,
    ghost_cells=1, device='cuda', sweeps_per_launch=10)``: n x n nodes on the
    unit square, dx = 1/(n-1), dt = min(0.5 dx / U, 0.25 dx^2 / nu) by
    default.  The Jacobi sweeps (and their transposes) run ``sweeps_per_launch``
    (at most 48) per launch on LDS-resident tiles
    (pde_jacobi_neumann0_multisweep_f32; 2 or fewer keeps the two-sweep
    row-tile kernels); the values do not depend on it."""

    def __init__(self, n, nu=0.01, U=1.0, dt=None, jacobi_sweeps=50, ghost_cells=1,
                 device="cuda", sweeps_per_launch=10):
        self.n, self.g = int(n), int(ghost_cells)
        self.dx = 1.0 / (self.n - 1)
        self.nu, self.U = float(nu), float(U)
        self.dt = float(dt) if dt is not None else min(0.5 * self.dx / self.U,
                                                       0.25 * self.dx ** 2 / self.nu)
        self.sweeps = int(jacobi_sweeps)
        self.sweeps_per_launch = max(1, int(sweeps_per_launch))
        if self.sweeps_per_launch > MAX_SWEEPS_PER_LAUNCH:
            raise ValueError(f"sweeps_per_launch {self.sweeps_per_launch} exceeds "
                             f"{MAX_SWEEPS_PER_LAUNCH}")
        self._multi = self.sweeps_per_launch > 2
        self._btabs = {}
        self.device = torch.device(device)
        P = self.n + 2 * self.g
        self.shape = (P, P, 1)
        z = lambda *e: torch.zeros(self.shape + e, dtype=torch.float32, device=self.device)
        self.u_bc = LegacyGridBoundary(z(2), self.dx, self.g)
        self.u_bc.dirichlet_BC("ALL", 0.0)
        self.u_bc.dirichlet_BC("+Y", float(self.U), 0)
        self.p_bc = LegacyGridBoundary(z(1), self.dx, self.g)
        self.p_bc.vonNeumann_BC("ALL", 0.0)
        self.wl = laplacian_weights(self.dx)
        self.wc = central_weights(self.dx)
        self.h2 = np.float32(self.dx * self.dx)
        self.inv = np.float32(0.25)
        # work buffers (the rhs ghosts stay 0: no kernel writes them)
        self.b = {k: z(*e) for k, e in (("u2", (2,)), ("us2", (2,)), ("rhs", (1,)),
                                         ("pf", (1,)), ("p0", (1,)))}
        m = np.zeros(self.shape + (1,), np.float32)
        m[self.g:P - self.g, self.g:P - self.g] = 1.0
        self.mask = torch.from_numpy(m).to(self.device)

    # --------------------------------------------------------------- forward
    def check_fields(self, vel=(), pres=(), lead=()):
        """Velocity fields must be ``lead + (X, Y, 1, 2)`` and pressure fields
        ``lead + (X, Y, 1, 1)`` of this grid (the kernels take their extents
        from the grid, not from the tensors), contiguous FP32 on the GPU."""
        for t, c in [(t, 2) for t in vel] + [(t, 1) for t in pres]:
            N.check_field(t, "field")
            want = tuple(lead) + self.shape + (c,)
            if tuple(t.shape) != want:
                raise AssertionError(f"field of shape {tuple(t.shape)}; expected {want}")

    def step(self, u, p, u_out, p_out):
        """One projection step from (u, p) into (u_out, p_out); u_out / p_out may
        be u / p.  All work on the current stream, no allocation."""
        self.check_fields((u, u_out), (p, p_out))
        b = self.b
        self._predict(u, b["us2"], b["rhs"], u_scratch=u is u_out)
        self._sweeps(p, b["rhs"], p_out, b["pf"], b["p0"], 1)
        self._correct(p_out, b["us2"], u_out)
        return u_out, p_out

    def _tables(self, batch):
        """(velocity, pressure) boundary tables for ``batch`` instances stored
        one after another: the grid's own, or for a batch their ops repeated
        per instance (cached).  The grid's tables must exist."""
        if batch == 1:
            return self.u_bc.table, self.p_bc.table
        if batch not in self._btabs:
            n = self.shape[0] * self.shape[1] * self.shape[2]
            self._btabs[batch] = (_repeat_table(self.u_bc.table, batch, 2 * n, self.device),
                                  _repeat_table(self.p_bc.table, batch, n, self.device))
        return self._btabs[batch]

    def _predict(self, u, us2, rhs, u_scratch=False, batch=1, u2buf=None):
        """u -> the boundary-fixed predictor us2 and the pressure source rhs
        (``batch`` instances one after another, one launch per stage).  With
        u_scratch (u is the step's output, overwritten at its end) the velocity
        boundary is applied to u in place instead of into a copy (u2buf, the
        grid's own work buffer by default)."""
        g, s = self.g, N.stream_of(u)
        grid = self.shape
        self.u_bc._ensure_table(u[0] if u.dim() == 5 else u)     # one instance's shape
        ut, _ = self._tables(batch)
        if u_scratch:
            u2 = ut.apply_inplace(u, None)
        else:
            u2 = ut.apply(u, self.b["u2"] if u2buf is None else u2buf, None)
        # grad, convection product, diffusion, F and the Euler update in one
        # launch (the values of the five separate kernels), then the boundary
        # on that result in place
        N.call("pde_cavity_predict_batch_f32", u2.data_ptr(), us2.data_ptr(), batch, *grid, g,
               *self.wl, *self.wc, np.float32(self.nu), np.float32(self.dt), s)
        ut.apply_inplace(us2, None)
        N.call("pde_divergence_batch_f32", us2.data_ptr(), rhs.data_ptr(), batch, *grid, 1, g,
               *self.wc, np.float32(1.0 / self.dt), s)

    def _sweep_plan(self):
        """Sweeps per launch, in order.  With several per launch the
        remainder goes first, so that the transposed launches (last to
        first) start with a full multi-sweep one."""
        k = self.sweeps_per_launch
        if not self._multi:
            return [2] * (self.sweeps // 2) + [1] * (self.sweeps % 2)
        return ([self.sweeps % k] if self.sweeps % k else []) + [k] * (self.sweeps // k)

    def _sweeps(self, p, rhs, p_out, pf, p0, batch):
        """The Jacobi sweeps from p into p_out (``batch`` instances stored one
        after another).  Neumann-0 walls are folded into each sweep
        (pde_jacobi_neumann0_f32: the same values as applying p_bc, then
        sweeping).  With sweeps_per_launch k > 2, k sweeps per launch on
        LDS-resident tiles (pde_jacobi_neumann0_multisweep_f32) and the
        remainder in one more; otherwise two per launch
        (pde_jacobi_neumann0_x2_f32) and a single one when the count is odd.
        Every form gives the same values.  Launches alternate between p_out
        and pf so that none writes what it reads and the last one writes
        p_out."""
        s = N.stream_of(p)
        plan = self._sweep_plan()
        cur = p
        for j, m in enumerate(plan):
            dst = p_out if (len(plan) - 1 - j) % 2 == 0 else pf
            if dst is cur:                       # p_out is p on the first launch
                N.call("pde_copy_f32", cur.data_ptr(), p0.data_ptr(), cur.numel(), s)
                cur = p0
            if self._multi and m > 1:
                N.call("pde_jacobi_neumann0_multisweep_f32", cur.data_ptr(), rhs.data_ptr(),
                       dst.data_ptr(), batch, *self.shape, self.g, self.h2, self.inv, m, s)
            elif m == 2:
                N.call("pde_jacobi_neumann0_x2_f32", cur.data_ptr(), rhs.data_ptr(),
                       dst.data_ptr(), batch, *self.shape, self.g, self.h2, self.inv, s)
            elif batch == 1:
                N.call("pde_jacobi_neumann0_f32", cur.data_ptr(), rhs.data_ptr(), dst.data_ptr(),
                       *self.shape, self.g, self.h2, self.inv, s)
            else:
                N.call("pde_jacobi_neumann0_batch_f32", cur.data_ptr(), rhs.data_ptr(),
                       dst.data_ptr(), batch, *self.shape, self.g, self.h2, self.inv, s)
            cur = dst
        if cur is not p_out:
            N.call("pde_copy_f32", cur.data_ptr(), p_out.data_ptr(), p_out.numel(), s)

    def _correct(self, p_out, us2, u_out, batch=1):
        """u_out = us2 - dt * grad(B_p(p_out))."""
        # B_p (Neumann 0 on every face) folded into the gradient: one launch
        # with the values of boundary + grad + Euler
        N.call("pde_cavity_correct_batch_f32", p_out.data_ptr(), us2.data_ptr(),
               u_out.data_ptr(), batch, *self.shape, self.g, *self.wc, np.float32(self.dt),
               N.stream_of(p_out))

    def run(self, u0, p0, steps, use_graph=True):
        """``steps`` steps from (u0, p0); returns the final (u, p)."""
        u, p = u0.clone(), p0.clone()
        if steps <= 0:
            return u, p
        self.step(u, p, u, p)
        if steps == 1:
            return u, p
        if not use_graph:
            for _ in range(steps - 1):
                self.step(u, p, u, p)
            return u, p
        torch.cuda.synchronize(self.device)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            self.step(u, p, u, p)
        for _ in range(steps - 1):
            graph.replay()
        return u, p

    def kinetic_energy(self, u):
        return float(0.5 * self.dx * self.dx * (u * u * self.mask).sum())

    # ---------------------------------------------------------- differentiable
    def diff_step(self, u, p):
        """The step as differentiable functions (fresh outputs, zero ghosts)."""
        u2 = A.boundary(u, self.u_bc)
        du = A.grad(u2, self.dx, self.g)
        conv = _MatVec.apply(du, u2)
        diff = A.laplacian(u2, self.dx, self.g, self.nu)
        F = _LinComb.apply((-1.0, 1.0), conv, diff)
        us = A.forward_euler(u2, F, self.dt)
        us2 = A.boundary(us, self.u_bc)
        rhs = A.divergence(us2.unsqueeze(3), self.dx, self.g, 1.0 / self.dt)
        for _ in range(self.sweeps):
            p = _Jacobi.apply(A.boundary(p, self.p_bc), rhs, self.g, self.h2, self.inv)
        gp = A.grad(A.boundary(p, self.p_bc), self.dx, self.g).squeeze(3)
        return A.forward_euler(us2, gp, -self.dt), p

    # ------------------------------------------------------------ transpose
    def _adjoint_buffers(self):
        if getattr(self, "ga", None) is None:
            z = lambda *e: torch.zeros(self.shape + e, dtype=torch.float32, device=self.device)
            self.ga = {k: z(*e) for k, e in (("gpf", (1,)), ("gpa", (1,)), ("gpb", (1,)),
                                             ("grhs", (1,)), ("gus2", (2,)))}
        return self.ga

    def prepare_adjoint(self, like_u, like_p):
        """Boundary tables and their transposes on the device (call before
        capturing ``step_adjoint`` in a graph)."""
        self.u_bc._ensure_table(like_u)
        self.p_bc._ensure_table(like_p)
        self.u_bc.table.prepare_adjoint()
        self.p_bc.table.prepare_adjoint()
        self._adjoint_buffers()

    def step_adjoint(self, u, gu_out, gp_out, gu, gp):
        """Transpose of ``step`` at the state u (the step's input velocity):
        from the gradients (gu_out, gp_out) of the step's outputs, the
        gradients (gu, gp) of its inputs (u, p).  gu / gp must not alias
        gu_out / gp_out.  Preallocated buffers only, current stream."""
        self.check_fields((u, gu_out, gu), (gp_out, gp))
        a = self._adjoint_buffers()
        self.prepare_adjoint(u, gp)
        first, spare = self._sweep_buffers(gp, a["gpb"], a["gpa"])
        gpf = self._adj_correct(gu_out)
        self._adj_sweeps(first, spare, a["grhs"], 1, start=(gpf, gp_out))   # ends in gp
        self._adj_predict(u, gu_out, a["grhs"], gu)
        return gu, gp

    def _sweep_launches(self):
        return len(self._sweep_plan())

    def _sweep_buffers(self, gp, b0, b1):
        """(first, spare) for the transposed sweeps so that the last launch
        writes gp: the launches alternate first -> spare -> first ..."""
        return (gp, b1) if self._sweep_launches() % 2 == 0 else (b0, gp)

    def _adj_correct(self, gu_out, batch=1, gpf=None):
        """Transpose of the correction u_out = us2 - dt * grad(B_p(p_out)) in
        p_out, up to the gradient p_out also receives from the step's output:
        gpf = B_p^T(grad^T(-dt * gu_out)) (gpf: work buffer, the grid's own by
        default), returned."""
        g, s = self.g, N.stream_of(gu_out)
        gpf = self.ga["gpf"] if gpf is None else gpf
        N.call("pde_cavity_adjoint_correct_batch_f32", gu_out.data_ptr(), gpf.data_ptr(), batch,
               *self.shape, g, *self.wc, np.float32(self.dt), s)
        self._tables(batch)[1].adjoint_inplace(gpf)
        return gpf

    def _adj_sweeps(self, g, spare, grhs, batch, start=None):
        """The sweeps p <- J(B_p(p), rhs), transposed last to first, from g
        (``batch`` instances one after another), launched as the forward's
        launches are; rhs's gradient gathers in grhs from 0 (the first
        multi-sweep launch starts it, else it is cleared first; its ghosts
        stay 0).  With start = (gpf, gp_out) the sweeps start from gpf + gp_out
        (pde_lincomb_f32's values), summed by the first multi-sweep launch as it
        loads them, or into g first.  Returns the buffer (g or spare) holding
        p's gradient."""
        s = N.stream_of(g)
        cur, nxt = g, spare
        plan = self._sweep_plan()
        fold = start is not None and self._multi and plan and plan[-1] > 1
        if start is not None and not fold:
            gpf, gp_out = start
            N.call("pde_lincomb_f32", gpf.data_ptr(), gp_out.data_ptr(), None, None,
                   np.float32(1.0), np.float32(1.0), np.float32(0), np.float32(0), 2,
                   g.data_ptr(), g.numel(), s)
        if self._multi:
            # the forward's launches transposed, last to first (each launch's
            # sweeps are one linear map; transposes compose in reverse)
            if not plan or plan[-1] == 1:       # no multi-sweep launch comes first
                grhs.zero_()
            for j, m in enumerate(reversed(plan)):
                if m == 1:
                    name = ("pde_jacobi_neumann0_adjoint_f32" if batch == 1 else
                            "pde_jacobi_neumann0_adjoint_batch_f32")
                    extra = () if batch == 1 else (batch,)
                    N.call(name, cur.data_ptr(), nxt.data_ptr(), grhs.data_ptr(), *extra,
                           *self.shape, self.g, self.h2, self.inv, s)
                else:
                    src, src2 = (start if fold and j == 0 else (cur, None))
                    N.call("pde_jacobi_neumann0_adjoint_multisweep_f32", src.data_ptr(),
                           src2.data_ptr() if src2 is not None else None, nxt.data_ptr(),
                           grhs.data_ptr(), batch, *self.shape, self.g, self.h2, self.inv, m,
                           int(j > 0 or plan[-1] == 1), s)
                cur, nxt = nxt, cur
            return cur
        grhs.zero_()
        for _ in range(self.sweeps // 2):          # two transposed sweeps per launch
            N.call("pde_jacobi_neumann0_adjoint_x2_f32", cur.data_ptr(), nxt.data_ptr(),
                   grhs.data_ptr(), batch, *self.shape, self.g, self.h2, self.inv, s)
            cur, nxt = nxt, cur
        for _ in range(self.sweeps % 2):
            if batch == 1:
                N.call("pde_jacobi_neumann0_adjoint_f32", cur.data_ptr(), nxt.data_ptr(),
                       grhs.data_ptr(), *self.shape, self.g, self.h2, self.inv, s)
            else:
                N.call("pde_jacobi_neumann0_adjoint_batch_f32", cur.data_ptr(), nxt.data_ptr(),
                       grhs.data_ptr(), batch, *self.shape, self.g, self.h2, self.inv, s)
            cur, nxt = nxt, cur
        return cur

    def _adj_predict(self, u, gu_out, grhs, gu, batch=1, u2buf=None, gus2=None):
        """Transpose of the predictor and the correction in us2: from rhs's
        gradient grhs and gu_out, u's gradient gu (u is the step's input;
        u2buf and gus2: work buffers, the grid's own by default)."""
        g, s = self.g, N.stream_of(u)
        grid = self.shape
        u2buf = self.b["u2"] if u2buf is None else u2buf
        gus2 = self.ga["gus2"] if gus2 is None else gus2
        ut, _ = self._tables(batch)
        # the forward value this transpose reads: u2 = B_u(u)
        ut.apply(u, u2buf, None)
        # rhs = div(us2) / dt; u_out = us2 - ...; us2 = B_u(us)
        N.call("pde_cavity_adjoint_div_batch_f32", grhs.data_ptr(), gu_out.data_ptr(),
               gus2.data_ptr(), batch, *grid, g, *self.wc, np.float32(1.0 / self.dt), s)
        ut.adjoint_inplace(gus2)                                  # now the gradient of us
        # us = u2 + dt * F, F = -(grad u2) u2 + nu lap(u2): the transposes of
        # the Euler update, F, the diffusion, the convection product and grad
        # in one launch (the values of the separate transposed kernels)
        N.call("pde_cavity_adjoint_predict_batch_f32", gus2.data_ptr(), u2buf.data_ptr(),
               gu.data_ptr(), batch, *grid, g, *self.wl, *self.wc, np.float32(self.nu),
               np.float32(self.dt), s)
        # u2 = B_u(u)
        ut.adjoint_inplace(gu)

    def kinetic_energy_rollout(self, u0, p0, steps, checkpoint_every=10, use_graph=False):
        """J(u_N) = 1/2 sum |u_N|^2 dx^2 (interior), differentiable in u0 and
        p0 through ``CavityRollout`` (``checkpoint_every=None``: every step's
        velocity kept, no recompute)."""
        ro = CavityRollout(self, steps, checkpoint_every, use_graph)
        uN, _ = _CavityRolloutFn.apply(u0, p0, ro)
        return _KineticEnergy.apply(uN, self.mask, float(self.dx * self.dx))

    def kinetic_energy_rollout_autograd(self, u0, p0, steps, checkpoint_every=10):
        """The same J through torch autograd over ``diff_step`` (segments of
        ``checkpoint_every`` steps recomputed in the backward pass)."""
        self.u_bc._ensure_table(u0)
        self.p_bc._ensure_table(p0)

        def segment(u, p, L):
            for _ in range(L):
                u, p = self.diff_step(u, p)
            return u, p

        u, p = u0, p0
        done = 0
        while done < steps:
            L = min(checkpoint_every, steps - done)
            u, p = cp.checkpoint(segment, u, p, L, use_reentrant=False)
            done += L
        return _KineticEnergy.apply(u, self.mask, float(self.dx * self.dx))


class CavityRollout:
    """``steps`` projection steps and their transpose on preallocated buffers.

    ``cav`` is a ProjectionCavity or a ProjectionCavityBatch (fields with a
    leading batch axis).  ``checkpoint_every=None`` keeps the input velocity
    of every step (steps fields; the transpose needs nothing else); K > 0
    keeps (u, p) every K steps and recomputes a segment's velocities before
    its transposed steps.  With ``use_graph`` the forward step and the
    transposed step are each captured once in a HIP graph (on fixed buffers)
    and replayed for every step.  One object holds the states of its last
    ``forward_states`` call: ``backward_states`` transposes that rollout."""

    def __init__(self, cav, steps, checkpoint_every=None, use_graph=False):
        self.cav = cav
        self.steps = int(steps)
        self.store_all = checkpoint_every is None
        self.K = self.steps if self.store_all else max(1, min(int(checkpoint_every), self.steps))
        self.use_graph = bool(use_graph)
        self._graphs = {}

    def _prepare(self, u0, p0):
        """Buffers, kept (with the graphs that point at them) while the shapes
        and device stay the same."""
        cav = self.cav
        cav.prepare_adjoint(u0, p0)
        key = (tuple(u0.shape), tuple(p0.shape), u0.device)
        if getattr(self, "_key", None) == key:
            return
        self._key = key
        self._graphs = {}
        E = lambda like, *n: torch.empty(n + tuple(like.shape), dtype=like.dtype,
                                         device=like.device)
        self.U, self.P = E(u0), E(p0)
        self.AU = E(u0)                                    # the transposed step's state
        self.gU, self.gP, self.gU2, self.gP2 = E(u0), E(p0), E(u0), E(p0)
        nseg = (self.steps + self.K - 1) // self.K
        if self.store_all:
            self.hist = E(u0, max(self.steps, 1))
        else:
            self.hist = E(u0, self.K)
            self.ck_u, self.ck_p = E(u0, nseg), E(p0, nseg)

    def _fwd(self):
        self._run("fwd", lambda: self.cav.step(self.U, self.P, self.U, self.P))

    def _adj(self):
        """One transposed step from the gradient pair of parity self._par into
        the other pair (two graphs, one per direction, no copies)."""
        src = (self.gU, self.gP) if self._par == 0 else (self.gU2, self.gP2)
        dst = (self.gU2, self.gP2) if self._par == 0 else (self.gU, self.gP)
        self._run(("adj", self._par),
                  lambda: self.cav.step_adjoint(self.AU, src[0], src[1], dst[0], dst[1]))
        self._par ^= 1

    def _run(self, key, fn):
        if not self.use_graph:
            fn()
            return
        gr = self._graphs.get(key)
        if gr is None:
            torch.cuda.synchronize(self.U.device)
            gr = torch.cuda.CUDAGraph()
            with torch.cuda.graph(gr):
                fn()
            self._graphs[key] = gr
        gr.replay()

    def forward_states(self, u0, p0):
        N.check_field(u0, "u0")
        N.check_field(p0, "p0")
        self._prepare(u0, p0)
        self.U.copy_(u0)
        self.P.copy_(p0)
        for n in range(self.steps):
            if self.store_all:
                self.hist[n].copy_(self.U)
            elif n % self.K == 0:
                self.ck_u[n // self.K].copy_(self.U)
                self.ck_p[n // self.K].copy_(self.P)
            self._fwd()
        return self.U.clone(), self.P.clone()

    def backward_states(self, guN, gpN):
        if getattr(self, "_key", None) is None:
            raise RuntimeError("backward_states needs a forward_states call first")
        for g, ref in ((guN, self.U), (gpN, self.P)):
            N.check_field(g, "gradient")
            if g.shape != ref.shape:
                raise AssertionError(f"gradient of shape {tuple(g.shape)}; the rollout's "
                                     f"state is {tuple(ref.shape)}")
        self.gU.copy_(guN)
        self.gP.copy_(gpN)
        self._par = 0
        nseg = (self.steps + self.K - 1) // self.K
        for sg in reversed(range(nseg)):
            n0 = sg * self.K
            L = min(self.K, self.steps - n0)
            if not self.store_all:
                self.U.copy_(self.ck_u[sg])
                self.P.copy_(self.ck_p[sg])
                for i in range(L):
                    self.hist[i].copy_(self.U)
                    self._fwd()
            for i in reversed(range(L)):
                self.AU.copy_(self.hist[n0 + i] if self.store_all else self.hist[i])
                self._adj()
        gu, gp = (self.gU, self.gP) if self._par == 0 else (self.gU2, self.gP2)
        return gu.clone(), gp.clone()


class _CavityRolloutFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, p0, ro):
        uN, pN = ro.forward_states(u0.detach().contiguous(), p0.detach().contiguous())
        ctx.ro = ro
        return uN, pN

    @staticmethod
    def backward(ctx, guN, gpN):
        ro = ctx.ro
        guN = torch.zeros_like(ro.U) if guN is None else guN.contiguous()
        gpN = torch.zeros_like(ro.P) if gpN is None else gpN.contiguous()
        gu0, gp0 = ctx.ro.backward_states(guN, gpN)
        return gu0, gp0, None


class ProjectionCavityBatch:
    """``batch`` cavities of one grid and one set of parameters stepped
    together (one per seed): u (batch, X, Y, 1, 2) and p (batch, X, Y, 1, 1)
    with the instances one after another.  Every stage runs as one launch for
    all instances: the predictor, divergence, Jacobi sweeps (several per
    launch, as in ``ProjectionCavity``), correction, and the boundaries (the
    grid's tables repeated per instance).  Each instance gets the same bits
    as ``ProjectionCavity.step`` alone."""

    def __init__(self, n, batch, **kw):
        self.cav = ProjectionCavity(n, **kw)
        self.batch = int(batch)
        if self.batch < 1:
            raise ValueError("batch must be >= 1")
        z = lambda *e: torch.zeros((self.batch,) + self.cav.shape + e, dtype=torch.float32,
                                   device=self.cav.device)
        self.us2, self.rhs, self.pf, self.p0, self.u2 = z(2), z(1), z(1), z(1), z(2)
        # the boundary tables, built here (never inside a graph capture)
        self.cav.u_bc._ensure_table(self.us2[0])
        self.cav.p_bc._ensure_table(self.rhs[0])
        self.cav._tables(self.batch)

    def step(self, u, p, u_out, p_out):
        """One step of every instance from (u, p) into (u_out, p_out) (which
        may be u / p)."""
        cav, B = self.cav, self.batch
        cav.check_fields((u, u_out), (p, p_out), (B,))
        cav._predict(u, self.us2, self.rhs, u_scratch=u is u_out, batch=B, u2buf=self.u2)
        cav._sweeps(p, self.rhs, p_out, self.pf, self.p0, B)
        cav._correct(p_out, self.us2, u_out, batch=B)
        return u_out, p_out

    def prepare_adjoint(self, like_u, like_p):
        self.cav.prepare_adjoint(like_u[0], like_p[0])
        for t in self.cav._tables(self.batch):
            t.prepare_adjoint()
        if getattr(self, "gpb", None) is None:
            z = lambda *e: torch.zeros((self.batch,) + self.cav.shape + e, dtype=torch.float32,
                                       device=self.cav.device)
            self.gpb, self.gpa, self.grhs, self.gpf, self.gus2 = z(1), z(1), z(1), z(1), z(2)

    def step_adjoint(self, u, gu_out, gp_out, gu, gp):
        """Transpose of ``step`` for every instance (each stage one launch for
        all instances); same arguments as ``ProjectionCavity.step_adjoint``
        with a leading batch axis."""
        cav, B = self.cav, self.batch
        cav.check_fields((u, gu_out, gu), (gp_out, gp), (B,))
        self.prepare_adjoint(u, gp)
        first, spare = cav._sweep_buffers(gp, self.gpb, self.gpa)
        gpf = cav._adj_correct(gu_out, batch=B, gpf=self.gpf)
        cav._adj_sweeps(first, spare, self.grhs, B, start=(gpf, gp_out))     # ends in gp
        cav._adj_predict(u, gu_out, self.grhs, gu, batch=B, u2buf=self.u2, gus2=self.gus2)
        return gu, gp

    def kinetic_energy_rollout(self, u0, p0, steps, checkpoint_every=10, use_graph=False):
        """Sum over the instances of J_i = 1/2 sum |u_i,N|^2 dx^2 (interior),
        differentiable in u0 and p0 (each instance's gradient is its own)."""
        ro = CavityRollout(self, steps, checkpoint_every, use_graph)
        uN, _ = _CavityRolloutFn.apply(u0, p0, ro)
        return _KineticEnergy.apply(uN, self.cav.mask, float(self.cav.dx * self.cav.dx))

    def run(self, u0, p0, steps, use_graph=True):
        """``steps`` steps of every instance; returns the final (u, p)."""
        u, p = u0.clone(), p0.clone()
        if steps <= 0:
            return u, p
        self.step(u, p, u, p)
        if steps == 1:
            return u, p
        if not use_graph:
            for _ in range(steps - 1):
                self.step(u, p, u, p)
            return u, p
        torch.cuda.synchronize(self.cav.device)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            self.step(u, p, u, p)
        for _ in range(steps - 1):
            graph.replay()
        return u, p
