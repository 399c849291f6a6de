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


GRAPH_CHUNK = 50   # steps per captured graph when every state is kept


def _copy(src, dst):
    """Device copy through the library's kernel (safe inside graph capture)."""
    N.call("pde_copy_f32", src.data_ptr(), dst.data_ptr(), src.numel(), N.stream_of(src))


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
        from .time_step.fused import HeatStep
        self._hs = HeatStep(boundary, dx, ghost_cells, alpha, dt)

    def _step(self, u, out):
        """u -> boundary in place -> out = fused heat step(u) (one launch on
        2D scalar fields whose boundary table allows it)."""
        self._hs.step_into(u, out)

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
                # hist[i] becomes u2 at step n0 + i
                self._step(hist[i], hist[i + 1] if i + 1 < L else tmp)
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
    alpha=1, src_idx=None, src_wavelet=None, rec_idx=None, use_graph=False)``.

    ``rollout(c2, u0, v0)`` returns ``(u_N, v_N, records)`` with ``records`` of
    shape (steps, n_receivers) (u after each step at the receivers' flat
    indices).  Per step n: boundary in place on u; ``v[src] += dt *
    src_wavelet[n]``; ``v_next = v + dt*(c2*(alpha*lap(u2)))``; ``u_next = u2 +
    dt*(v if n == 0 else v_next)``; sample u_next.  Differentiable in c2, u0
    and v0.

    Every per-step input that changes with n (the source row, the receiver
    row, the first-step flag) is read through a device-resident step counter,
    so with ``use_graph=True`` one HIP graph per loop segment (forward,
    recompute, adjoint) is captured once and replayed for every segment; the
    eager path runs the same kernel sequence.

    ``reverse=True`` (2D scalar fields, checkpointed): the backward pass does
    not recompute segments.  The forward also records the scalars the
    boundary overwrites at each step (one float per boundary op and step), and
    every segment but the first is walked backwards from the checkpoint at its
    end, two transposed steps per launch, rebuilding each state from the next
    one (``pde_wave_reverse_adjoint_pair_f32``).  The gradients in u0 and v0
    are those of the recompute path bit for bit; the c2 gradient differs only
    by the rounding of the rebuilt states (checkpoints bound how far that
    rounding can travel).  Segments of odd length use the recompute path, and
    so do steps 0 and 1 of the first segment (the plain-Euler step 0 is not
    reversible); the rest of the first segment is walked back by pairs."""

    def __init__(self, boundary, dx, ghost_cells, dt, steps, checkpoint_every=50, alpha=1.0,
                 src_idx=None, src_wavelet=None, rec_idx=None, use_graph=False, reverse=False):
        self.bc = boundary
        self.w = laplacian_weights(dx)
        self.g = int(ghost_cells)
        self.alpha = np.float32(alpha)
        self.dt = np.float32(dt)
        self.steps = int(steps)
        # checkpoint_every=None keeps every step's state (no recompute in the
        # backward; steps + 1 fields of memory)
        self.store_all = checkpoint_every is None
        self.K = GRAPH_CHUNK if self.store_all else max(1, min(int(checkpoint_every), self.steps))
        self.K = min(self.K, self.steps)
        self.src_idx, self.src_wavelet, self.rec_idx = src_idx, src_wavelet, rec_idx
        if src_idx is not None:
            if src_wavelet is None or src_wavelet.dim() != 2 or src_wavelet.shape[0] < self.steps \
                    or src_wavelet.shape[1] != src_idx.numel():
                raise ValueError("src_wavelet must be (steps, n_sources)")
            self.src_wavelet = src_wavelet.contiguous()
        self.use_graph = bool(use_graph)
        self.reverse = bool(reverse)
        if self.reverse and self.store_all:
            raise ValueError("reverse=True walks back from checkpoints: give checkpoint_every")
        self._saving = False
        self._shape = None
        self._c2 = None
        # the boundary table the buffers, ring plan and graphs were built for;
        # held so that its device arrays (which captured graphs point at) live
        # as long as they are used
        self._table = None
        self._graphs = {}

    # ---------------------------------------------------------------- buffers
    def _prepare(self, like, c2):
        # device tables must exist before any graph capture (no uploads inside)
        self.bc._ensure_table(like)
        table = self.bc.table
        table.prepare_adjoint()
        shape = tuple(like.shape)
        if self._shape == shape and self._dev == like.device and \
                (self._c2 is None) == (c2 is None) and self._table is table:
            return
        # a new boundary table (set_BC, shift_group, ...) invalidates the
        # captured graphs and the reversed walk's ring plan and save buffer
        self._graphs = {}
        self._table = table
        dev = like.device
        self._shape, self._dev = shape, dev
        E = lambda *s: torch.empty(s + shape, dtype=like.dtype, device=dev)
        self._a, self._b, self._v = E(), E(), E()
        self._hist = E((self.steps if self.store_all else self.K) + 1)
        self._vt = E()
        self._lu, self._lv, self._gu2, self._gvn = E(), E(), E(), E()
        self._gc2 = E()
        self._c2 = E() if c2 is not None else None
        nseg = (self.steps + self.K - 1) // self.K
        if self.store_all:
            self._cu = self._cv = None
        else:
            # reverse: one more checkpoint, the final state
            self._cu, self._cv = E(nseg + self.reverse), E(nseg + self.reverse)
        nrec = self.rec_idx.numel() if self.rec_idx is not None else 0
        self._nrec = nrec
        self._rec = torch.empty((self.steps, max(nrec, 1)), dtype=like.dtype, device=dev)
        self._grec = torch.zeros_like(self._rec)
        # two step counters, used alternately by the fused 2D steps (slot 0
        # only by the generic path)
        self._ctr = torch.zeros(2, dtype=torch.int64, device=dev)
        self._fused = sum(1 for s_ in shape[:3] if s_ > 1) == 2 and shape[3] == 1
        numel = int(np.prod(shape))
        for name, idx in (("src_idx", self.src_idx), ("rec_idx", self.rec_idx)):
            if idx is not None and (idx.dtype != torch.int64 or idx.device != dev
                                    or not idx.is_contiguous() or idx.dim() != 1):
                raise TypeError(f"{name} must be a contiguous 1-D int64 tensor on {dev}")
            if idx is not None and idx.numel():
                lo, hi = int(idx.min()), int(idx.max())
                if lo < 0 or hi >= numel:
                    raise IndexError(f"{name} holds {lo}..{hi}, outside a field of {numel} "
                                     "values")
        if self._fused:
            self._build_point_sets(shape)
        if self.reverse:
            self._prepare_reverse(shape, E)
        self._graphs = {}

    def _prepare_reverse(self, shape, E):
        from .ring_plan import DeviceRingPlan, build_ring_plan
        table = self.bc.table
        if not self._fused:
            raise ValueError("reverse=True needs a 2D scalar field")
        if table.nops and table.chunk_start is None:
            raise ValueError("reverse=True needs a boundary table that chunks")
        p1, p2 = [int(s_) for s_ in shape[:3] if s_ > 1]
        self._ring = DeviceRingPlan(build_ring_plan(table.host_ops, p1, p2, self.g), self._dev)
        self._nops = int(table.nops)
        self._save = torch.empty((self.steps, max(self._nops, 1)), dtype=torch.float32,
                                 device=self._dev)
        self._ru, self._rv = [E(), E()], [E(), E()]

    def _build_point_sets(self, shape):
        """Row-indexed source / receiver lists of the fused 2D steps (struct
        pde_point_set): points sorted by padded row, list order kept within a
        row."""
        p1, p2 = [int(s_) for s_ in shape[:3] if s_ > 1]
        dev = self._dev
        self._pts_keep = []

        def build(idx, data, stride, scale):
            e = idx.detach().cpu().numpy().astype(np.int64)
            if e.size and (e.min() < 0 or e.max() >= p1 * p2):
                raise IndexError("point index outside the field")
            rows, cols = e // p2, e % p2
            order = np.argsort(rows, kind="stable")
            row_start = np.searchsorted(rows[order], np.arange(p1 + 1)).astype(np.int32)
            arrs = [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in
                    (row_start, cols[order].astype(np.int32), order.astype(np.int32))]
            self._pts_keep += arrs + [data]
            rec = np.zeros(1, N.POINT_SET_DTYPE)
            rec[0] = (arrs[0].data_ptr(), arrs[1].data_ptr(), arrs[2].data_ptr(),
                      data.data_ptr(), stride, np.float32(scale), 0)
            self._pts_keep.append(rec)
            return rec

        self._ps_src = (build(self.src_idx, self.src_wavelet, self.src_wavelet.shape[1], self.dt)
                        if self.src_idx is not None else None)
        if self._nrec:
            self._ps_rec = build(self.rec_idx, self._rec, self._rec.shape[1], 1.0)
            self._ps_grec = build(self.rec_idx, self._grec, self._grec.shape[1], 1.0)
        else:
            self._ps_rec = self._ps_grec = None

    @staticmethod
    def _ps(rec):
        return rec.ctypes.data if rec is not None else None

    def _set_counter(self, n):
        self._ctr[0].fill_(int(n))

    def _c(self, slot):
        return self._ctr.data_ptr() + 8 * slot

    # ---------------------------------------------------------- step pieces
    def _step(self, u, v, c2, out, record):
        st = N.stream_of(u)
        self.bc.apply_inplace(u)
        if self.src_idx is not None:
            N.call("pde_scatter_add_step_f32", v.data_ptr(), self.src_idx.data_ptr(),
                   self.src_wavelet.data_ptr(), self.src_idx.numel(), self.src_idx.numel(),
                   self.dt, self._ctr.data_ptr(), st)
        N.call("pde_wave_step_ctr_f32", u.data_ptr(), v.data_ptr(),
               c2.data_ptr() if c2 is not None else None, out.data_ptr(), v.data_ptr(),
               *_grid(u), self.g, *self.w, self.alpha, self.dt, self._ctr.data_ptr(), st)
        if record and self._nrec:
            N.call("pde_gather_step_f32", out.data_ptr(), self.rec_idx.data_ptr(),
                   self._rec.data_ptr(), self._nrec, self._rec.shape[1], self._ctr.data_ptr(), st)
        N.call("pde_counter_add", self._ctr.data_ptr(), 1, st)

    def _step_fused(self, u, v, c2, out, record, slot):
        """One launch per step after the boundary (2D scalar fields): source,
        u/v update, receivers and the counter move slot -> 1 - slot."""
        if self._saving and self._nops:
            self.bc.table.apply_inplace_save(u, self.bc._time_tensor(u), self._save, self._c(slot))
        else:
            self.bc.apply_inplace(u)
        N.call("pde_wave_step_fused_f32", u.data_ptr(), v.data_ptr(),
               c2.data_ptr() if c2 is not None else None, out.data_ptr(), v.data_ptr(),
               *_grid(u), self.g, *self.w, self.alpha, self.dt, self._c(slot),
               self._c(1 - slot), self._ps(self._ps_src),
               self._ps(self._ps_rec) if record else None, N.stream_of(u))

    def _run_steps(self, pairs, c2, record):
        """Steps over (u, out) pairs with v = the segment's velocity buffer;
        the counter starts and ends in slot 0."""
        if not self._fused:
            for u, v, out in pairs:
                self._step(u, v, c2, out, record)
            return
        slot = 0
        for u, v, out in pairs:
            self._step_fused(u, v, c2, out, record, slot)
            slot = 1 - slot
        if slot:
            N.call("pde_counter_copy", self._c(1), self._c(0), 0, N.stream_of(self._ctr))

    def _fwd_segment(self, L, c2):
        """L steps from self._a / self._v; the state ends in self._a."""
        a, b = self._a, self._b
        pairs = []
        for _ in range(L):
            pairs.append((a, self._v, b))
            a, b = b, a
        self._run_steps(pairs, c2, True)
        if a is not self._a:
            _copy(a, self._a)

    def _hist_segment(self, i0, L, c2, record):
        """L steps from hist[i0] / vt writing hist[i0 + i + 1]; afterwards
        hist[i0 + i] holds u2 of that step."""
        self._run_steps([(self._hist[i], self._vt, self._hist[i + 1])
                         for i in range(i0, i0 + L)], c2, record)

    def _adj_segment(self, i0, L, c2, need_c2):
        """L transposed steps on hist[i0 + L - 1] .. hist[i0], last to first.
        The gradients live in lu / lv on entry and on exit."""
        st = N.stream_of(self._lu)
        lu, lv, gu2, gvn = self._lu, self._lv, self._gu2, self._gvn
        if self._fused:
            # counter n0 + L in slot 0 -> n0 + L - 1 in slot 1, then down by one
            # per step, alternating slots
            N.call("pde_counter_copy", self._c(0), self._c(1), -1, st)
            slot = 1
            for i in reversed(range(i0, i0 + L)):
                N.call("pde_wave_step_adjoint_fused_f32", lu.data_ptr(), lv.data_ptr(),
                       c2.data_ptr() if c2 is not None else None, self._hist[i].data_ptr(),
                       gu2.data_ptr(), gvn.data_ptr(),
                       self._gc2.data_ptr() if need_c2 else None, *_grid(lu), self.g, *self.w,
                       self.alpha, self.dt, self._c(slot), self._c(1 - slot),
                       self._ps(self._ps_grec), st)
                slot = 1 - slot
                self.bc.table.adjoint_inplace(gu2)
                lu, gu2 = gu2, lu
                lv, gvn = gvn, lv
            if lu is not self._lu:
                _copy(lu, self._lu)
                _copy(lv, self._lv)
            return
        N.call("pde_counter_add", self._ctr.data_ptr(), -1, st)
        for i in reversed(range(i0, i0 + L)):
            if self._nrec:
                N.call("pde_scatter_add_step_f32", lu.data_ptr(), self.rec_idx.data_ptr(),
                       self._grec.data_ptr(), self._nrec, self._grec.shape[1], np.float32(1.0),
                       self._ctr.data_ptr(), st)
            N.call("pde_wave_step_adjoint_ctr_f32", lu.data_ptr(), lv.data_ptr(),
                   c2.data_ptr() if c2 is not None else None, self._hist[i].data_ptr(),
                   gu2.data_ptr(), gvn.data_ptr(),
                   self._gc2.data_ptr() if need_c2 else None, *_grid(lu), self.g, *self.w,
                   self.alpha, self.dt, self._ctr.data_ptr(), st)
            self.bc.table.adjoint_inplace(gu2)
            lu, gu2 = gu2, lu
            lv, gvn = gvn, lv
            N.call("pde_counter_add", self._ctr.data_ptr(), -1, st)
        if lu is not self._lu:
            _copy(lu, self._lu)
            _copy(lv, self._lv)

    def _rev_segment(self, L, c2, need_c2):
        """L / 2 pair launches walking the states back from (ru[0], rv[0]),
        the state at the segment's end, with the counter (slot 0) at the
        segment's last step.  The gradients live in lu / lv on entry and on
        exit."""
        st = N.stream_of(self._lu)
        lu, lv, gu2, gvn = self._lu, self._lv, self._gu2, self._gvn
        ua, va, ub, vb = self._ru[0], self._rv[0], self._ru[1], self._rv[1]
        slot = 0
        for _ in range(L // 2):
            N.call("pde_wave_reverse_adjoint_pair_f32", ua.data_ptr(), va.data_ptr(),
                   lu.data_ptr(), lv.data_ptr(), c2.data_ptr() if c2 is not None else None,
                   ub.data_ptr(), vb.data_ptr(), gu2.data_ptr(), gvn.data_ptr(),
                   self._gc2.data_ptr() if need_c2 else None, *_grid(lu), self.g, *self.w,
                   self.alpha, self.dt, self._c(slot), self._c(1 - slot),
                   self._ps(self._ps_src), self._ps(self._ps_grec), self._save.data_ptr(),
                   self._nops, self._ring.ptr, st)
            slot = 1 - slot
            ua, ub, va, vb = ub, ua, vb, va
            lu, gu2, lv, gvn = gu2, lu, gvn, lv
        if lu is not self._lu:
            _copy(lu, self._lu)
            _copy(lv, self._lv)

    def _run(self, key, fn, *args):
        """Run fn eagerly, or through a captured graph replayed per call."""
        if not self.use_graph:
            fn(*args)
            return
        gr = self._graphs.get(key)
        if gr is None:
            torch.cuda.synchronize(self._dev)
            gr = torch.cuda.CUDAGraph()
            with torch.cuda.graph(gr):
                fn(*args)
            self._graphs[key] = gr
        gr.replay()

    # ---------------------------------------------------------------- passes
    def forward_states(self, c2, u0, v0):
        self._prepare(u0, c2)
        if c2 is not None:
            self._c2.copy_(c2)
            c2 = self._c2
        self._set_counter(0)
        nseg = (self.steps + self.K - 1) // self.K
        if self.store_all:
            self._hist[0].copy_(u0)
            self._vt.copy_(v0)
            for s in range(nseg):
                n0 = s * self.K
                L = min(self.K, self.steps - n0)
                self._run(("fwd_all", s, c2 is None), self._hist_segment, n0, L, c2, True)
            uN, vN = self._hist[self.steps], self._vt
        else:
            self._a.copy_(u0)
            self._v.copy_(v0)
            self._saving = self.reverse
            try:
                for s in range(nseg):
                    L = min(self.K, self.steps - s * self.K)
                    self._cu[s].copy_(self._a)
                    self._cv[s].copy_(self._v)
                    self._run(("fwd", L, c2 is None, self._saving), self._fwd_segment, L, c2)
            finally:
                self._saving = False
            if self.reverse:
                self._cu[nseg].copy_(self._a)
                self._cv[nseg].copy_(self._v)
            uN, vN = self._a, self._v
        rec = self._rec[:, :self._nrec] if self._nrec else self._rec[:, :0]
        return uN.clone(), vN.clone(), rec.clone()

    def backward_states(self, c2, gu, gv, grec, need_c2):
        if self.bc.table is not self._table:
            raise RuntimeError("the boundary conditions changed between the rollout's forward "
                               "and backward passes")
        self.bc.table.prepare_adjoint()
        if c2 is not None:
            self._c2.copy_(c2)
            c2 = self._c2
        self._lu.copy_(gu)
        self._lv.copy_(gv)
        if self._nrec:
            if grec is None:
                self._grec.zero_()
            else:
                self._grec[:, :self._nrec].copy_(grec)
        if need_c2:
            self._gc2.zero_()
        nseg = (self.steps + self.K - 1) // self.K
        if self.store_all:
            for s in reversed(range(nseg)):
                n0 = s * self.K
                L = min(self.K, self.steps - n0)
                self._set_counter(n0 + L)
                self._run(("adj_all", s, c2 is None, need_c2), self._adj_segment, n0, L, c2,
                          need_c2)
        else:
            for s in reversed(range(nseg)):
                n0 = s * self.K
                L = min(self.K, self.steps - n0)
                # steps walked back by pair launches: the whole segment, or in
                # the first one every step but 0 and 1 (pairs rebuild states
                # down to step 2; step 0 is the plain-Euler one, which is not
                # reversible), recomputed below from the checkpoint
                R = (L if s > 0 else L - 2) if self.reverse and L % 2 == 0 else 0
                if R > 0:
                    self._ru[0].copy_(self._cu[s + 1])
                    self._rv[0].copy_(self._cv[s + 1])
                    self._set_counter(n0 + L - 1)
                    self._run(("rev", R, c2 is None, need_c2), self._rev_segment, R, c2,
                              need_c2)
                if R == L:
                    continue
                Lr = L - R
                self._hist[0].copy_(self._cu[s])
                self._vt.copy_(self._cv[s])
                self._set_counter(n0)
                self._run(("rec", Lr, c2 is None), self._hist_segment, 0, Lr, c2, False)
                self._run(("adj", Lr, c2 is None, need_c2), self._adj_segment, 0, Lr, c2,
                          need_c2)
        return (self._gc2.clone() if need_c2 else None), self._lu.clone(), self._lv.clone()

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
        for name, t in (("v0", v0), ("c2", c2)):
            if t is not None and t.shape != u0.shape:
                raise AssertionError(f"{name} has shape {tuple(t.shape)}, u0 {tuple(u0.shape)}")
        c2d = c2.detach() if c2 is not None else None
        uN, vN, rec = ro.forward_states(c2d, u0.detach(), v0.detach())
        ctx.ro, ctx.c2 = ro, c2d
        ctx.need_c2 = c2 is not None and ctx.needs_input_grad[0]
        ctx.like = u0
        return uN, vN, rec

    @staticmethod
    def backward(ctx, guN, gvN, grec):
        ref = ctx.like
        guN = torch.zeros_like(ref) if guN is None else guN.contiguous()
        gvN = torch.zeros_like(ref) if gvN is None else gvN.contiguous()
        grec = grec.contiguous() if grec is not None and grec.numel() else None
        gc2, gu0, gv0 = ctx.ro.backward_states(ctx.c2, guN, gvN, grec, ctx.need_c2)
        return gc2, gu0, gv0, None
