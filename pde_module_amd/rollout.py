"""Differentiable explicit time loops with checkpointing.

``HeatRollout`` runs the fused heat step (boundary in place, then
``u_next = u2 + dt * alpha * lap(u2)``) for ``steps`` steps; ``WaveRollout``
runs the fused u/v wave step of the reference's standing-wave loop
(``examples/wave2d/standingWave.py:70-76``: plain Euler on the first step,
symplectic afterwards) with a per-node c², an optional point source added to v
before each step and receivers sampling u after each step.

Both are ``torch.autograd.Function``s.  The forward pass keeps only a
checkpoint of the state every ``checkpoint_every`` steps.  The backward pass
walks the segments from last to first: it recomputes the segment from its
checkpoint, writing each step's state straight into a segment history (the
boundary is then applied to it in place, so the history holds exactly the
boundary-applied states the forward steps read), and runs the fused transpose
kernels backwards through it.  Recomputation is deterministic, so the gradient
does not depend on ``checkpoint_every``.

Memory: (steps / checkpoint_every) checkpoints plus one segment history of
``checkpoint_every`` fields; ``checkpoint_every = steps`` keeps everything.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _native as N
from .FDM.module.laplacian import laplacian_weights


def _grid(x):
    return tuple(int(s) for s in x.shape[:3])


class HeatRollout:
    """``HeatRollout(boundary, dx, ghost_cells, alpha, dt, steps, checkpoint_every)``;
    ``rollout(u0)`` returns u after ``steps`` steps, differentiable in u0."""

    def __init__(self, boundary, dx, ghost_cells, alpha, dt, steps, checkpoint_every=50):
        self.bc = boundary
        self.w = laplacian_weights(dx)
        self.g = int(ghost_cells)
        self.alpha = np.float32(alpha)
        self.dt = np.float32(dt)
        self.steps = int(steps)
        self.K = max(1, int(checkpoint_every))

    def _step(self, u, out):
        """u -> boundary in place -> out = fused heat step(u)."""
        self.bc.apply_inplace(u)
        N.call("pde_heat_step_f32", u.data_ptr(), out.data_ptr(), *_grid(u), int(u.shape[3]),
               self.g, *self.w, self.alpha, self.dt, N.stream_of(u))

    def forward_states(self, u0):
        """Runs the loop; returns (u_N, checkpoints {n: u^n})."""
        a, b = u0.clone(), torch.empty_like(u0)
        ckpt = {0: u0.clone()}
        for n in range(self.steps):
            self._step(a, b)
            a, b = b, a
            if (n + 1) % self.K == 0 and n + 1 < self.steps:
                ckpt[n + 1] = a.clone()
        return a, ckpt

    def backward_states(self, ckpt, g_final):
        """Adjoint of the loop: the gradient with respect to u0."""
        shape = g_final.shape
        lam = g_final.clone()
        gu2 = torch.empty_like(lam)
        starts = sorted(ckpt)
        hist = None
        for s_i in reversed(range(len(starts))):
            n0 = starts[s_i]
            n1 = starts[s_i + 1] if s_i + 1 < len(starts) else self.steps
            L = n1 - n0
            if hist is None or hist.shape[0] < L:
                hist = torch.empty((L,) + tuple(shape), dtype=lam.dtype, device=lam.device)
            tmp = torch.empty_like(lam)
            hist[0].copy_(ckpt[n0])
            for i in range(L):
                self.bc.apply_inplace(hist[i])            # hist[i] = u2 at step n0 + i
                nxt = hist[i + 1] if i + 1 < L else tmp
                N.call("pde_heat_step_f32", hist[i].data_ptr(), nxt.data_ptr(), *_grid(lam),
                       int(shape[3]), self.g, *self.w, self.alpha, self.dt, N.stream_of(lam))
            for i in reversed(range(L)):
                N.call("pde_heat_step_adjoint_f32", lam.data_ptr(), gu2.data_ptr(), *_grid(lam),
                       int(shape[3]), self.g, *self.w, self.alpha, self.dt, N.stream_of(lam))
                self.bc.table.adjoint(gu2, lam)
        return lam

    def __call__(self, u0):
        self.bc._ensure_table(u0)
        return _HeatFn.apply(u0, self)


class _HeatFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, ro):
        N.check_field(u0, "u0")
        uN, ckpt = ro.forward_states(u0.detach())
        ctx.ro, ctx.ckpt = ro, ckpt
        return uN

    @staticmethod
    def backward(ctx, g):
        g = g.contiguous() if g is not None else None
        return ctx.ro.backward_states(ctx.ckpt, g), None


class WaveRollout:
    """``WaveRollout(boundary, dx, ghost_cells, dt, steps, checkpoint_every,
    alpha=1, src_idx=None, src_wavelet=None, rec_idx=None)``.

    ``rollout(c2, u0, v0)`` returns ``(u_N, v_N, records)`` with ``records`` of
    shape (steps, n_receivers) (u after each step at the receivers' flat
    indices).  Per step n: boundary in place on u; ``v[src] += dt *
    src_wavelet[n]``; ``v_next = v + dt*(c2*(alpha*lap(u2)))``; ``u_next = u2 +
    dt*(v if n == 0 else v_next)``; sample u_next.  Differentiable in c2, u0
    and v0."""

    def __init__(self, boundary, dx, ghost_cells, dt, steps, checkpoint_every=50, alpha=1.0,
                 src_idx=None, src_wavelet=None, rec_idx=None):
        self.bc = boundary
        self.w = laplacian_weights(dx)
        self.g = int(ghost_cells)
        self.alpha = np.float32(alpha)
        self.dt = np.float32(dt)
        self.steps = int(steps)
        self.K = max(1, int(checkpoint_every))
        self.src_idx, self.src_wavelet, self.rec_idx = src_idx, src_wavelet, rec_idx
        if src_idx is not None and (src_wavelet is None or src_wavelet.shape[0] < self.steps):
            raise ValueError("src_wavelet needs one row per step")

    def _step(self, u, v, c2, out, n, rec):
        st = N.stream_of(u)
        self.bc.apply_inplace(u)
        if self.src_idx is not None:
            N.call("pde_scatter_add_f32", v.data_ptr(), self.src_idx.data_ptr(),
                   self.src_wavelet[n].data_ptr(), self.src_idx.numel(), self.dt, st)
        N.call("pde_wave_step_f32", u.data_ptr(), v.data_ptr(),
               c2.data_ptr() if c2 is not None else None, out.data_ptr(), v.data_ptr(),
               *_grid(u), self.g, *self.w, self.alpha, self.dt, 1 if n == 0 else 0, st)
        if rec is not None:
            N.call("pde_gather_f32", out.data_ptr(), self.rec_idx.data_ptr(), rec[n].data_ptr(),
                   self.rec_idx.numel(), st)

    def forward_states(self, c2, u0, v0):
        a, b, v = u0.clone(), torch.empty_like(u0), v0.clone()
        nrec = self.rec_idx.numel() if self.rec_idx is not None else 0
        rec = torch.empty((self.steps, max(nrec, 1)), dtype=u0.dtype, device=u0.device) if nrec else None
        ckpt = {0: (u0.clone(), v0.clone())}
        for n in range(self.steps):
            self._step(a, v, c2, b, n, rec)
            a, b = b, a
            if (n + 1) % self.K == 0 and n + 1 < self.steps:
                ckpt[n + 1] = (a.clone(), v.clone())
        if rec is None:
            rec = torch.zeros((self.steps, 0), dtype=u0.dtype, device=u0.device)
        return a, v, rec, ckpt

    def backward_states(self, c2, ckpt, gu, gv, grec, need_c2):
        st = N.stream_of(gu)
        shape = gu.shape
        lu, lv = gu.clone(), gv.clone()
        gu2, gvn = torch.empty_like(lu), torch.empty_like(lu)
        gc2 = torch.zeros_like(lu) if need_c2 else None
        starts = sorted(ckpt)
        hist = None
        vtmp = torch.empty_like(lu)
        utmp = torch.empty_like(lu)
        for s_i in reversed(range(len(starts))):
            n0 = starts[s_i]
            n1 = starts[s_i + 1] if s_i + 1 < len(starts) else self.steps
            L = n1 - n0
            if hist is None or hist.shape[0] < L:
                hist = torch.empty((L,) + tuple(shape), dtype=lu.dtype, device=lu.device)
            u_c, v_c = ckpt[n0]
            hist[0].copy_(u_c)
            vtmp.copy_(v_c)
            for i in range(L):
                n = n0 + i
                nxt = hist[i + 1] if i + 1 < L else utmp
                self._step(hist[i], vtmp, c2, nxt, n, None)
            # after the recompute hist[i] holds u2 of step n0 + i
            for i in reversed(range(L)):
                n = n0 + i
                if grec is not None and self.rec_idx is not None:
                    N.call("pde_scatter_add_f32", lu.data_ptr(), self.rec_idx.data_ptr(),
                           grec[n].data_ptr(), self.rec_idx.numel(), np.float32(1.0), st)
                N.call("pde_wave_step_adjoint_f32", lu.data_ptr(), lv.data_ptr(),
                       c2.data_ptr() if c2 is not None else None, hist[i].data_ptr(),
                       gu2.data_ptr(), gvn.data_ptr(), gc2.data_ptr() if gc2 is not None else None,
                       *_grid(lu), self.g, *self.w, self.alpha, self.dt, 1 if n == 0 else 0, st)
                self.bc.table.adjoint(gu2, lu)
                lv, gvn = gvn, lv
        return gc2, lu, lv

    def __call__(self, c2, u0, v0):
        self.bc._ensure_table(u0)
        return _WaveFn.apply(c2, u0, v0, self)


class _WaveFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, c2, u0, v0, ro):
        N.check_field(u0, "u0")
        N.check_field(v0, "v0")
        if c2 is not None:
            N.check_field(c2, "c2")
        uN, vN, rec, ckpt = ro.forward_states(c2.detach() if c2 is not None else None,
                                              u0.detach(), v0.detach())
        ctx.ro, ctx.ckpt = ro, ckpt
        ctx.c2 = c2.detach() if c2 is not None else None
        ctx.need_c2 = c2 is not None and ctx.needs_input_grad[0]
        return uN, vN, rec

    @staticmethod
    def backward(ctx, guN, gvN, grec):
        ref = ctx.ckpt[0][0]
        guN = torch.zeros_like(ref) if guN is None else guN.contiguous()
        gvN = torch.zeros_like(ref) if gvN is None else gvN.contiguous()
        grec = grec.contiguous() if grec is not None and grec.numel() else None
        gc2, gu0, gv0 = ctx.ro.backward_states(ctx.c2, ctx.ckpt, guN, gvN, grec, ctx.need_c2)
        return gc2, gu0, gv0, None
