"""Host resolution of boundary groups into the device op table (struct
pde_bc_op of include/pde_hip.h).

The reference applies a boundary as ``out = copy(in)`` followed by writes whose
values depend only on ``in`` (and t): the functional GridBoundary launches one
kernel per group in insertion order (``functional_boundary.py:278-291``, kernel
``:356-400``), the legacy one launches one table kernel over (node, component)
(``pm/deprecated/gridBoundary.py:231-242``, kernel ``:267-302``).  So the final
value of each written scalar is the formula of its *last* writer applied to
``in``.  These functions list every write in launch order, keep the last one per
destination, and drop writes that restore ``in[dst]`` (the copy already does),
so one launch over the result reproduces the reference exactly.
"""
from __future__ import annotations

import numpy as np

from ..._native import BC_COPY, BC_GHOST_D, BC_GHOST_N, BC_OP_DTYPE, BC_SET
from .flags import DIRICHLET, VON_NEUMANN


def _strides(grid_shape):
    X, Y, Z = grid_shape
    return np.array([Y * Z, Z, 1], dtype=np.int64)


def _neumann_coef(s, dx32):
    """-sign(s) * 2 * dx in FP32 (exact: a power-of-two scaling of dx)."""
    return np.where(s > 0, np.float32(-2.0), np.float32(2.0)).astype(np.float32) * np.float32(dx32)


class _Writes:
    """Writes in the order a serial run of the reference's launches makes
    them.  ``key`` orders writes inside one launch: the reference runs those in
    parallel, and only ill-posed tables (e.g. a shifted legacy group landing on
    another boundary node) make two of them hit one scalar; such collisions are
    resolved as if the threads ran in index order."""

    def __init__(self):
        self.parts = []
        self.keys = []
        self.launch = 0

    def next_launch(self):
        self.launch += 1

    def add(self, dst, src, kind, value, coef, tmax, key):
        n = len(dst)
        if n == 0:
            return
        rec = np.zeros(n, dtype=BC_OP_DTYPE)
        rec["dst"] = dst
        rec["src"] = src
        rec["kind"] = kind
        rec["value"] = value
        rec["coef"] = coef
        rec["tmax"] = tmax
        self.parts.append(rec)
        k = np.empty((n, 2), dtype=np.int64)
        k[:, 0] = self.launch
        k[:, 1] = key
        self.keys.append(k)

    def resolve(self):
        if not self.parts:
            return np.zeros(0, dtype=BC_OP_DTYPE)
        allw = np.concatenate(self.parts)
        keys = np.concatenate(self.keys)
        order = np.lexsort((keys[:, 1], keys[:, 0]))  # stable: launch, then key
        rev = allw[order][::-1]
        _, first = np.unique(rev["dst"], return_index=True)  # last writer per dst
        ops = rev[first]
        keep = ~((ops["kind"] == BC_COPY) & (ops["src"] == ops["dst"]))
        return np.ascontiguousarray(ops[keep])


def functional_ops(grid_shape, ncomp, ghost_cells, groups, dx32):
    """``groups``: ordered iterable of (ids, boundary_type, output_ids, value)
    with value a ConstantValue/RampValue."""
    g = ghost_cells
    sn = _strides(grid_shape)
    Y, Z = grid_shape[1], grid_shape[2]
    elig = [a for a in range(3) if grid_shape[a] > 1]
    W = _Writes()
    for ids, btype, output_ids, value in groups:
        W.next_launch()
        ids = np.asarray(ids, dtype=np.int64).ravel()
        pos = np.arange(len(ids), dtype=np.int64)
        comps = [int(c) for c in (output_ids if isinstance(output_ids, (list, tuple)) else [output_ids])]
        v32, t32 = value.value32, value.tmax32
        if btype not in (DIRICHLET, VON_NEUMANN):
            raise ValueError(f"unknown boundary type {btype}")
        # per node: the node write (phase c), then for each axis a and each
        # selected component m the ghost write (phase ncomp + a*len(comps) + m)
        nph = ncomp + 3 * max(len(comps), 1)
        if btype == DIRICHLET:
            # out[node] = val: the whole vector, selected components replaced.
            for c in range(ncomp):
                dst = ids * ncomp + c
                if c in comps:
                    W.add(dst, -1, BC_SET, v32, 0.0, t32, pos * nph + c)
                else:
                    W.add(dst, dst, BC_COPY, 0.0, 0.0, 0.0, pos * nph + c)
        idx = np.stack([ids // (Y * Z), (ids // Z) % Y, ids % Z])
        for a in elig:
            s = np.where(idx[a] == g, 1, np.where(idx[a] == grid_shape[a] - 1 - g, -1, 0))
            sel = s != 0
            q, sq, pq = ids[sel], s[sel].astype(np.int64), pos[sel]
            ghost, adj = q - sq * sn[a], q + sq * sn[a]
            for m, c in enumerate(comps):
                key = pq * nph + ncomp + a * len(comps) + m
                if btype == DIRICHLET:
                    W.add(ghost * ncomp + c, adj * ncomp + c, BC_GHOST_D, v32, 0.0, t32, key)
                else:
                    W.add(ghost * ncomp + c, adj * ncomp + c, BC_GHOST_N, v32,
                          _neumann_coef(sq, dx32), t32, key)
    return W.resolve()


def legacy_ops(grid_shape, ncomp, ijk, interior, btype, bvalue, dx32):
    """Legacy table: ijk/interior (nb, 3), btype/bvalue (nb, C)."""
    sn = _strides(grid_shape)
    node = np.asarray(ijk, dtype=np.int64) @ sn
    interior = np.asarray(interior)
    W = _Writes()
    # one launch; per entry b, component c, axis a: node write then ghost write
    for a in range(3):
        b = np.nonzero(interior[:, a] != 0)[0]
        s = interior[b, a].astype(np.int64)
        ghost, adj = node[b] - s * sn[a], node[b] + s * sn[a]
        coef = _neumann_coef(s, dx32)
        for c in range(ncomp):
            key = ((b.astype(np.int64) * ncomp + c) * 3 + a) * 2
            ty = np.asarray(btype)[b, c]
            v = np.asarray(bvalue, dtype=np.float32)[b, c]
            d = ty == DIRICHLET
            W.add(node[b][d] * ncomp + c, -1, BC_SET, v[d], 0.0, 0.0, key[d])
            W.add(ghost[d] * ncomp + c, adj[d] * ncomp + c, BC_GHOST_D, v[d], 0.0, 0.0, key[d] + 1)
            n = ty == VON_NEUMANN
            W.add(ghost[n] * ncomp + c, adj[n] * ncomp + c, BC_GHOST_N, v[n], coef[n], 0.0,
                  key[n] + 1)
    return W.resolve()


def transpose_table(ops):
    """Arrays of pde_bc_adjoint_f32 for a resolved op table: every written
    scalar is zeroed, then each distinct source gathers sign * g[dst] over the
    ops that read it (sorted by dst), plus g[src] itself unless it is written.
    Signs: COPY +1, Neumann ghost +1, Dirichlet ghost -1; SET reads nothing."""
    zero_idx = np.ascontiguousarray(ops["dst"], dtype=np.int64)
    reads = ops[ops["kind"] != BC_SET]
    sign = np.where(reads["kind"] == BC_GHOST_D, np.float32(-1.0), np.float32(1.0)).astype(np.float32)
    order = np.lexsort((reads["dst"], reads["src"]))
    src_sorted = reads["src"][order]
    cdst = np.ascontiguousarray(reads["dst"][order], dtype=np.int64)
    csign = np.ascontiguousarray(sign[order])
    src_idx, start = np.unique(src_sorted, return_index=True)
    off = np.append(start, len(src_sorted)).astype(np.int64)
    flags = np.isin(src_idx, zero_idx).astype(np.int32)
    return dict(zero_idx=zero_idx, src_idx=np.ascontiguousarray(src_idx, dtype=np.int64),
                flags=flags, off=off, cdst=cdst, csign=csign)
