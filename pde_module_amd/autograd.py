"""Reverse-mode differentiation of the uniform-grid operators.

The reference has no autodiff (no ``wp.Tape`` anywhere in PDE_Module).  Each
function here runs a forward kernel of libpdehip.so and, in ``backward``, its
exact transpose kernel.  The forward operators are viewed as linear maps from
the whole padded field to its interior nodes: their outputs here have zero
ghosts (the modules' outputs leave ghosts untouched), and the transposes write
the whole padded input gradient.

Gradients flow to the field arguments only; weights, ``alpha``, ``dt`` and the
boundary values are constants.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _native as N
from .FDM.module.grad import central_weights
from .FDM.module.laplacian import laplacian_weights


def _grid(x):
    return tuple(int(s) for s in x.shape[:3])


def _contig(g):
    return g if g.is_contiguous() else g.contiguous()


class _Laplacian(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, w, alpha, g):
        N.check_field(u, "u")
        out = torch.zeros_like(u)
        N.call("pde_laplacian_f32", u.data_ptr(), out.data_ptr(), *_grid(u), int(u.shape[3]), g,
               *w, np.float32(alpha), N.stream_of(u))
        ctx.meta = (w, alpha, g)
        return out

    @staticmethod
    def backward(ctx, gout):
        w, alpha, g = ctx.meta
        gout = _contig(gout)
        gin = torch.empty_like(gout)
        N.call("pde_laplacian_adjoint_f32", gout.data_ptr(), gin.data_ptr(), *_grid(gout),
               int(gout.shape[3]), g, *w, np.float32(alpha), N.stream_of(gout))
        return gin, None, None, None


class _Grad(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, w, alpha, g):
        N.check_field(u, "u")
        D = sum(1 for s in u.shape[:3] if s > 1)
        out = torch.zeros(tuple(u.shape[:4]) + (D,), dtype=u.dtype, device=u.device)
        N.call("pde_grad_f32", u.data_ptr(), out.data_ptr(), *_grid(u), int(u.shape[3]), g, *w,
               np.float32(alpha), N.stream_of(u))
        ctx.meta = (w, alpha, g)
        return out

    @staticmethod
    def backward(ctx, gout):
        w, alpha, g = ctx.meta
        gout = _contig(gout)
        gin = torch.empty(tuple(gout.shape[:4]), dtype=gout.dtype, device=gout.device)
        N.call("pde_grad_adjoint_f32", gout.data_ptr(), gin.data_ptr(), *_grid(gout),
               int(gout.shape[3]), g, *w, np.float32(alpha), N.stream_of(gout))
        return gin, None, None, None


class _Divergence(torch.autograd.Function):
    @staticmethod
    def forward(ctx, m, w, alpha, g):
        """m is (X, Y, Z, rows, D); returns (X, Y, Z, rows)."""
        N.check_field(m, "field")
        out = torch.zeros(tuple(m.shape[:4]), dtype=m.dtype, device=m.device)
        N.call("pde_divergence_f32", m.data_ptr(), out.data_ptr(), *_grid(m), int(m.shape[3]), g,
               *w, np.float32(alpha), N.stream_of(m))
        ctx.meta = (w, alpha, g, int(m.shape[4]))
        return out

    @staticmethod
    def backward(ctx, gout):
        w, alpha, g, D = ctx.meta
        gout = _contig(gout)
        gin = torch.empty(tuple(gout.shape[:4]) + (D,), dtype=gout.dtype, device=gout.device)
        N.call("pde_divergence_adjoint_f32", gout.data_ptr(), gin.data_ptr(), *_grid(gout),
               int(gout.shape[3]), g, *w, np.float32(alpha), N.stream_of(gout))
        return gin, None, None, None


class _Euler(torch.autograd.Function):
    @staticmethod
    def forward(ctx, y, k, dt):
        N.check_field(y, "y")
        N.check_field(k, "k")
        out = torch.empty_like(y)
        N.call("pde_forward_euler_f32", y.data_ptr(), k.data_ptr(), out.data_ptr(), y.numel(),
               np.float32(dt), N.stream_of(y))
        ctx.dt = dt
        return out

    @staticmethod
    def backward(ctx, g):
        g = _contig(g)
        gk = torch.empty_like(g)
        N.call("pde_scale_f32", g.data_ptr(), gk.data_ptr(), g.numel(), np.float32(ctx.dt),
               N.stream_of(g))
        return g, gk, None


class _Boundary(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, bc, t):
        bc._ensure_table(u)
        out = torch.empty_like(u)
        bc.table.apply(u, out, bc._time_tensor(u, t))
        ctx.bc = bc
        return out

    @staticmethod
    def backward(ctx, g):
        g = _contig(g)
        gin = torch.empty_like(g)
        ctx.bc.table.adjoint(g, gin)
        return gin, None, None


def laplacian(u, dx, ghost_cells, alpha=1.0):
    """Differentiable Laplacian of a (X, Y, Z, C) field (zero output ghosts)."""
    return _Laplacian.apply(u, laplacian_weights(dx), alpha, ghost_cells)


def grad(u, dx, ghost_cells, alpha=1.0):
    """Differentiable gradient: (X, Y, Z, N) -> (X, Y, Z, N, D)."""
    return _Grad.apply(u, central_weights(dx), alpha, ghost_cells)


def divergence(m, dx, ghost_cells, alpha=1.0):
    """Differentiable divergence: (X, Y, Z, rows, D) -> (X, Y, Z, rows); pass a
    vector field as (X, Y, Z, 1, D)."""
    return _Divergence.apply(m, central_weights(dx), alpha, ghost_cells)


def forward_euler(y, k, dt):
    """Differentiable y + dt * k."""
    return _Euler.apply(y, k, dt)


def boundary(u, bc, t=None):
    """Differentiable boundary application (gradient through the linear part of
    a functional or legacy GridBoundary built for u's shape)."""
    return _Boundary.apply(u, bc, t)
