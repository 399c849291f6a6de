"""Device side of a resolved boundary: the op table uploaded once at setup,
the 1-element time array and the staging buffer for in-place application.

The table is stored in conflict-closed chunk order (``tables.chunk_ops``), so
in-place application and its transpose are one full-chip launch each; when a
table cannot be chunked (a conflict component larger than a workgroup) the
two-launch gather/scatter forms are used."""
from __future__ import annotations

import numpy as np
import torch

from ... import _native as N


class DeviceBoundaryTable:
    def __init__(self, ops: np.ndarray, device: torch.device):
        from .tables import chunk_ops
        self.nops = int(len(ops))
        ch = chunk_ops(ops) if self.nops else None
        self.chunk_start = None
        if ch is not None:
            ops = ch[0]
            self.chunk_start = torch.from_numpy(ch[1]).to(device)
            self.nchunks = len(ch[1]) - 1
        self.needs_t = bool(self.nops and np.any((ops["tmax"] > 0)
                                                 | (ops["vkind"] != N.BCV_CONST)))
        # array-valued ops read t[1 + slot]: the time array passed must hold them
        arr = ops[ops["vkind"] == N.BCV_ARRAY] if self.nops else ops[:0]
        self.nslots = int(N.bc_slots(arr).max()) + 1 if len(arr) else 0
        raw = np.ascontiguousarray(ops).view(np.uint8) if self.nops else np.zeros(32, np.uint8)
        self.ops = torch.from_numpy(raw.copy()).to(device)
        self.staging = torch.empty(max(self.nops, 1), dtype=torch.float32, device=device)
        self.host_ops = ops
        self.device = device
        self._adj = None
        self._adj_chunks = None
        # Largest scalar index any op touches; checked against the field size at
        # every call so a bad table raises on the host instead of faulting.
        if self.nops:
            src = ops["src"][ops["kind"] != N.BC_SET]
            self.max_index = int(max(ops["dst"].max(), src.max() if len(src) else 0))
            self.min_index = int(min(ops["dst"].min(), src.min() if len(src) else 0))
        else:
            self.max_index, self.min_index = -1, 0

    def _check_bounds(self, numel: int) -> None:
        if self.min_index < 0 or self.max_index >= numel:
            raise IndexError(
                f"boundary table touches scalar index {self.min_index}..{self.max_index} "
                f"outside a field of {numel} values (node ids or shifted groups leave the grid)")

    def _check_t(self, t: torch.Tensor | None, u: torch.Tensor) -> None:
        if t is None:
            if self.nslots:
                raise AssertionError("a boundary with function values needs its value array")
            return
        if t.dtype != torch.float32 or t.device != u.device or not t.is_contiguous() \
                or t.numel() < 1 + self.nslots:
            raise AssertionError(f"the time array must be a contiguous float32 array of at "
                                 f"least {1 + self.nslots} values on the field's device")

    def apply(self, x: torch.Tensor, out: torch.Tensor, t: torch.Tensor | None) -> torch.Tensor:
        """out = copy(x) with the boundary writes (reads all from x)."""
        N.check_field(x, "input_array")
        self._check_t(t, x)
        N.check_field(out, "output_array")
        if out.shape != x.shape:
            raise AssertionError("boundary output must match the input's shape")
        self._check_bounds(x.numel())
        N.call("pde_bc_apply_f32", N.dev_ptr(x), N.dev_ptr(out), x.numel(), self.ops.data_ptr(),
               self.nops, t.data_ptr() if t is not None else None, N.stream_of(x))
        return out

    def apply_inplace(self, u: torch.Tensor, t: torch.Tensor | None) -> torch.Tensor:
        """The same writes applied to u itself (values gathered from the
        unmodified u first)."""
        N.check_field(u, "field")
        self._check_t(t, u)
        self._check_bounds(u.numel())
        if self.chunk_start is not None:
            N.call("pde_bc_apply_chunked_inplace_f32", N.dev_ptr(u), self.ops.data_ptr(),
                   self.chunk_start.data_ptr(), self.nchunks,
                   t.data_ptr() if t is not None else None, N.stream_of(u))
            return u
        N.call("pde_bc_apply_inplace_f32", N.dev_ptr(u), self.ops.data_ptr(), self.nops,
               t.data_ptr() if t is not None else None, self.staging.data_ptr(),
               N.stream_of(u))
        return u

    def apply_inplace_save(self, u: torch.Tensor, t: torch.Tensor | None, save: torch.Tensor,
                           step_ptr: int) -> torch.Tensor:
        """``apply_inplace`` that also records the value every written scalar
        held before the apply, op by op in table order, in row ``*step_ptr`` (a
        device int64) of ``save`` (rows of ``nops`` floats).  Chunked tables
        only; a table without ops writes nothing."""
        N.check_field(u, "field")
        self._check_t(t, u)
        self._check_bounds(u.numel())
        if not self.nops:
            return u
        if self.chunk_start is None:
            raise ValueError("recording the overwritten scalars needs a chunked boundary table")
        if save.dtype != torch.float32 or save.device != u.device or save.dim() != 2 \
                or save.shape[1] != self.nops or not save.is_contiguous():
            raise AssertionError(f"save must be a contiguous float32 (rows, {self.nops}) array "
                                 "on the field's device")
        N.call("pde_bc_apply_chunked_inplace_save_f32", N.dev_ptr(u), self.ops.data_ptr(),
               self.chunk_start.data_ptr(), self.nchunks, self.nops,
               t.data_ptr() if t is not None else None, save.data_ptr(), step_ptr,
               N.stream_of(u))
        return u

    def prepare_adjoint(self) -> None:
        """Uploads the transpose arrays (once); call before capturing a graph
        that applies the transpose."""
        if self._adj is not None:
            return
        if self.nslots:
            raise NotImplementedError(
                "a boundary value given by a Python function may read the field, so the "
                "boundary's transpose is not known: use constant, ramp or sine values in "
                "differentiated rollouts")
        if not self.nops:
            self._adj = (None, 0, None, None, None, None, None, 0)
            return
        from .tables import chunk_transpose, transpose_table
        T = transpose_table(self.host_ops)
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(self.device)
        C = chunk_transpose(T)
        self._adj_chunks = None
        if C is not None:
            self._adj_chunks = tuple(dev(C[k]) for k in ("src_idx", "flags", "off", "cdst",
                                                         "csign", "item_start", "zero_idx",
                                                         "zero_start")) + (len(C["item_start"]) - 1,)
        self._adj = (dev(T["zero_idx"]), len(T["zero_idx"]), dev(T["src_idx"]), dev(T["flags"]),
                     dev(T["off"]), dev(T["cdst"]), dev(T["csign"]), len(T["src_idx"]))
        self._adj_staging = torch.empty(max(len(T["src_idx"]), 1), dtype=torch.float32,
                                        device=self.device)

    def adjoint(self, gout: torch.Tensor, gin: torch.Tensor) -> torch.Tensor:
        """gin = transpose of the boundary's linear part applied to gout."""
        N.check_field(gout, "gout")
        N.check_field(gin, "gin")
        self._check_bounds(gout.numel())
        self.prepare_adjoint()
        z, nz, s, f, o, cd, cs, ns = self._adj
        p = lambda t: t.data_ptr() if t is not None else None
        N.call("pde_bc_adjoint_f32", N.dev_ptr(gout), N.dev_ptr(gin), gout.numel(), p(z), nz,
               p(s), p(f), p(o), p(cd), p(cs), ns, N.stream_of(gout))
        return gin

    def adjoint_inplace(self, g: torch.Tensor) -> torch.Tensor:
        """g <- transpose of the boundary's linear part applied to g (no copy)."""
        N.check_field(g, "g")
        self._check_bounds(g.numel())
        self.prepare_adjoint()
        z, nz, s, f, o, cd, cs, ns = self._adj
        if nz == 0 and ns == 0:
            return g
        p = lambda t: t.data_ptr() if t is not None else None
        if self._adj_chunks is not None:
            *arrs, nch = self._adj_chunks
            N.call("pde_bc_adjoint_chunked_inplace_f32", N.dev_ptr(g), *(a.data_ptr() for a in arrs),
                   nch, N.stream_of(g))
            return g
        N.call("pde_bc_adjoint_inplace_f32", N.dev_ptr(g), p(z), nz, p(s), p(f), p(o), p(cd),
               p(cs), ns, self._adj_staging.data_ptr(), N.stream_of(g))
        return g
