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

from ..._native import (BC_COPY, BC_GHOST_D, BC_GHOST_N, BC_OP_DTYPE, BC_SET, BCV_ARRAY,
                         BCV_CONST, bc_slot_bits)
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

    def add(self, dst, src, kind, value, coef, tmax, key, vkind=BCV_CONST):
        """value: FP32 values, or for vkind == BCV_ARRAY the slots (ints)."""
        n = len(dst)
        if n == 0:
            return
        rec = np.zeros(n, dtype=BC_OP_DTYPE)
        rec["dst"] = dst
        rec["src"] = src
        rec["kind"] = kind
        rec["vkind"] = vkind
        if vkind == BCV_ARRAY:
            rec["value"] = bc_slot_bits(np.broadcast_to(np.asarray(value, dtype=np.int64), (n,)))
        else:
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
    """``groups``: ordered iterable of (ids, boundary_type, output_ids, value,
    slot_base) with value a ConstantValue/RampValue/SineValue, or a
    FunctionValue whose vector for the group's node at position p is in slots
    slot_base + p*ncomp + c (``function_values.FunctionSlots``)."""
    g = ghost_cells
    sn = _strides(grid_shape)
    Y, Z = grid_shape[1], grid_shape[2]
    elig = [a for a in range(3) if grid_shape[a] > 1]
    W = _Writes()
    for ids, btype, output_ids, value, slot_base in groups:
        W.next_launch()
        ids = np.asarray(ids, dtype=np.int64).ravel()
        pos = np.arange(len(ids), dtype=np.int64)
        comps = [int(c) for c in (output_ids if isinstance(output_ids, (list, tuple)) else [output_ids])]
        vk = value.vkind
        t32 = value.tmax32
        # the value of component c for the nodes at positions p
        val = (lambda p, c: slot_base + p * ncomp + c) if vk == BCV_ARRAY else \
            (lambda p, c: value.value32)
        if btype not in (DIRICHLET, VON_NEUMANN):
            raise ValueError(f"unknown boundary type {btype}")
        # per node: the node write (phase c), then for each axis a and each
        # selected component m the ghost write (phase ncomp + a*len(comps) + m)
        nph = ncomp + 3 * max(len(comps), 1)
        if btype == DIRICHLET:
            # out[node] = val: the whole vector.  A record's vector is the
            # node's own with the selected components replaced
            # (boundary_functions.py:42-45); a function's is what it returns.
            for c in range(ncomp):
                dst = ids * ncomp + c
                if c in comps or vk == BCV_ARRAY:
                    W.add(dst, -1, BC_SET, val(pos, c), 0.0, t32, pos * nph + c, vk)
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
                    W.add(ghost * ncomp + c, adj * ncomp + c, BC_GHOST_D, val(pq, c), 0.0, t32,
                          key, vk)
                else:
                    W.add(ghost * ncomp + c, adj * ncomp + c, BC_GHOST_N, val(pq, c),
                          _neumann_coef(sq, dx32), t32, key, vk)
    return W.resolve()


def legacy_ops(grid_shape, ncomp, ijk, interior, btype, bvalue, dx32, func_groups=()):
    """Legacy table: ijk/interior (nb, 3), btype/bvalue (nb, C); then the
    function groups, each one launch after the table's in order
    (reference ``pm/deprecated/gridBoundary.py:244-245``, kernel
    ``pm/deprecated/func_boundary.py:96-134``): ``func_groups`` holds
    (entry ids, boundary_type, varID, value, slot_base) with value a
    RampValue/SineValue/ConstantValue record or a FunctionValue whose value for
    the entry at position p is in slot slot_base + p."""
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
    for ids, fbtype, var, value, slot_base in func_groups:
        # one thread per entry: for each axis with an interior direction,
        # Dirichlet writes the node and its ghost, Neumann the ghost
        W.next_launch()
        ids = np.asarray(ids, dtype=np.int64).ravel()
        pos = np.arange(len(ids), dtype=np.int64)
        var = int(var)
        vk, t32 = value.vkind, value.tmax32
        for a in range(3):
            sel = interior[ids, a] != 0
            b, p = ids[sel], pos[sel]
            s = interior[b, a].astype(np.int64)
            ghost, adj = node[b] - s * sn[a], node[b] + s * sn[a]
            vals = slot_base + p if vk == BCV_ARRAY else value.value32
            key = (p * 3 + a) * 2
            if fbtype == DIRICHLET:
                W.add(node[b] * ncomp + var, -1, BC_SET, vals, 0.0, t32, key, vk)
                W.add(ghost * ncomp + var, adj * ncomp + var, BC_GHOST_D, vals, 0.0, t32,
                      key + 1, vk)
            elif fbtype == VON_NEUMANN:
                W.add(ghost * ncomp + var, adj * ncomp + var, BC_GHOST_N, vals,
                      _neumann_coef(s, dx32), t32, key + 1, vk)
            else:
                raise ValueError(f"unknown boundary type {fbtype}")
    return W.resolve()


def transpose_table(ops):
    """Arrays of pde_bc_adjoint_f32 for a resolved op table: each distinct
    source gathers sign * g[dst] over the ops that read it (sorted by dst),
    plus g[src] itself unless it is written (flag 1); every written scalar
    that is not a source is zeroed (zero_idx: disjoint from src_idx, so the
    in-place form can zero and store in one launch).
    Signs: COPY +1, Neumann ghost +1, Dirichlet ghost -1; SET reads nothing."""
    dst = np.ascontiguousarray(ops["dst"], dtype=np.int64)
    reads = ops[ops["kind"] != BC_SET]
    sign = np.where(reads["kind"] == BC_GHOST_D, np.float32(-1.0), np.float32(1.0)).astype(np.float32)
    order = np.lexsort((reads["dst"], reads["src"]))
    src_sorted = reads["src"][order]
    cdst = np.ascontiguousarray(reads["dst"][order], dtype=np.int64)
    csign = np.ascontiguousarray(sign[order])
    src_idx, start = np.unique(src_sorted, return_index=True)
    off = np.append(start, len(src_sorted)).astype(np.int64)
    flags = np.isin(src_idx, dst).astype(np.int32)
    zero_idx = np.ascontiguousarray(dst[~np.isin(dst, src_idx)], dtype=np.int64)
    return dict(zero_idx=zero_idx, src_idx=np.ascontiguousarray(src_idx, dtype=np.int64),
                flags=flags, off=off, cdst=cdst, csign=csign)


# ----------------------------------------------------------------- chunking
# One full-chip launch can apply a table in place if it is split into
# conflict-closed chunks: no entry reads a scalar that an entry of another
# chunk writes.  Each workgroup then does its chunk's reads, a barrier and its
# writes.  Conflicts are rare (ops near corners whose source is itself a
# boundary node), so chunks are just runs of at most CHUNK entries cut at
# component boundaries.
CHUNK = 256


def _components(n, a, b):
    """Connected-component label per unit 0..n-1 for edges (a[i], b[i])."""
    lab = np.arange(n, dtype=np.int64)
    if len(a) == 0:
        return lab
    a = np.asarray(a, dtype=np.int64)
    b = np.asarray(b, dtype=np.int64)
    while True:
        la, lb = lab[a], lab[b]
        m = np.minimum(la, lb)
        before = lab.copy()
        np.minimum.at(lab, la, m)
        np.minimum.at(lab, lb, m)
        while True:                      # pointer jumping to the roots
            nl = lab[lab]
            if np.array_equal(nl, lab):
                break
            lab = nl
        if np.array_equal(lab, before):
            return lab


def _pack(labels, weight, cap):
    """Units grouped by component (stable), cut into runs whose total weight is
    at most cap without splitting a component.  Returns (order, run bounds in
    units) or None if a component alone exceeds cap."""
    order = np.argsort(labels, kind="stable")
    ls = labels[order]
    n = len(order)
    if n == 0:
        return order, np.zeros(1, np.int64)
    ends = np.append(np.flatnonzero(np.diff(ls)) + 1, n)          # component ends
    cw = np.concatenate([[0], np.cumsum(weight[order])])           # weight prefix
    comp_w = cw[ends] - cw[np.concatenate([[0], ends[:-1]])]
    if comp_w.max() > cap:
        return None
    bounds = [0]
    cur = 0
    wend = cw[ends]
    while cur < n:
        k = np.searchsorted(wend, cw[cur] + cap, side="right") - 1
        nxt = int(ends[k])
        if nxt <= cur:                   # zero-weight tail etc.
            nxt = int(ends[np.searchsorted(ends, cur, side="right")])
        bounds.append(nxt)
        cur = nxt
    return order, np.asarray(bounds, np.int64)


def chunk_ops(ops, cap=CHUNK):
    """(ops in chunk order, chunk_start) for pde_bc_apply_chunked_inplace_f32,
    or None when a conflict component is larger than a chunk."""
    n = len(ops)
    if n == 0:
        return ops, np.zeros(1, np.int64)
    dst = np.asarray(ops["dst"], np.int64)
    by_dst = np.argsort(dst, kind="stable")
    sd = dst[by_dst]
    reads = np.flatnonzero(ops["kind"] != BC_SET)
    src = np.asarray(ops["src"], np.int64)[reads]
    pos = np.minimum(np.searchsorted(sd, src), n - 1)
    hit = sd[pos] == src
    labels = _components(n, reads[hit], by_dst[pos[hit]])
    packed = _pack(labels, np.ones(n, np.int64), cap)
    if packed is None:
        return None
    order, bounds = packed
    return np.ascontiguousarray(ops[order]), bounds


def chunk_transpose(T, cap=CHUNK):
    """Chunk-ordered arrays of pde_bc_adjoint_chunked_inplace_f32 from a
    transpose_table dict (at most cap sources plus zero entries per chunk), or
    None when a conflict component is larger than that."""
    src, off = T["src_idx"], T["off"]
    zero = T["zero_idx"]
    ni, nz = len(src), len(zero)
    U = ni + nz
    if U == 0:
        return None
    loc = np.concatenate([src, zero]).astype(np.int64)           # scalar each unit writes
    by_loc = np.argsort(loc, kind="stable")
    sl = loc[by_loc]
    counts = np.diff(off)
    item_of_q = np.repeat(np.arange(ni, dtype=np.int64), counts)
    cd = np.asarray(T["cdst"], np.int64)
    pos = np.minimum(np.searchsorted(sl, cd), U - 1)
    hit = sl[pos] == cd
    labels = _components(U, item_of_q[hit], by_loc[pos[hit]])
    # zero entries count too, so no chunk collects a long tail of them
    weight = np.ones(U, np.int64)
    packed = _pack(labels, weight, cap)
    if packed is None:
        return None
    order, bounds = packed
    is_item = order < ni
    item_start, zero_start = [0], [0]
    items, zeros = [], []
    for c in range(len(bounds) - 1):
        seg = order[bounds[c]:bounds[c + 1]]
        m = seg < ni
        items.append(seg[m])
        zeros.append(seg[~m] - ni)
        item_start.append(item_start[-1] + int(m.sum()))
        zero_start.append(zero_start[-1] + int((~m).sum()))
    perm = np.concatenate(items) if items else np.zeros(0, np.int64)
    zperm = np.concatenate(zeros) if zeros else np.zeros(0, np.int64)
    assert is_item.sum() == len(perm) == ni
    cnt = counts[perm]
    new_off = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int64)
    # gather each source's (cdst, csign) segment in the new order
    seg_start = np.repeat(off[:-1][perm] - new_off[:-1], cnt)
    qidx = np.arange(int(new_off[-1]), dtype=np.int64) + seg_start
    return dict(src_idx=np.ascontiguousarray(src[perm], np.int64),
                flags=np.ascontiguousarray(T["flags"][perm], np.int32),
                off=new_off,
                cdst=np.ascontiguousarray(cd[qidx], np.int64),
                csign=np.ascontiguousarray(np.asarray(T["csign"])[qidx], np.float32),
                item_start=np.asarray(item_start, np.int64),
                zero_idx=np.ascontiguousarray(zero[zperm], np.int64),
                zero_start=np.asarray(zero_start, np.int64))
