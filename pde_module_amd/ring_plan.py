"""Boundary ring of a 2D scalar field for the time-reversed wave gradient
(``pde_wave_reverse_adjoint_pair_f32``).

The pair kernel runs nodes far from the edge as register row tiles that know
nothing of the boundary, and the outer ``depth`` rows and columns as pieces
staged in LDS that apply the boundary's saved scalars and its transpose
themselves.  This module builds the piece table on the host from the resolved
boundary op table (in the chunk order the device table uses, whose op index
is the save slot) and checks every assumption the kernel makes; a table that
breaks one raises ``ValueError`` (callers then keep the recompute path).

Depth: every node the boundary writes or reads, and every node its transpose
touches, lies within ``depth - 2`` of the edge, so a tile's step-n values
(one node beyond its output) never need a boundary fix-up.  At least 4 (the
tiles' first lane reaches 4 columns before the output) and ``ghost + 2``.
"""
from __future__ import annotations

import numpy as np

from . import _native as N
from .FDM.boundary.tables import transpose_table

PIECE = 32       # ring nodes along the edge per piece
RING_CAP = 448   # nodes of a piece grown by two (the kernel's LDS arrays)
MAX_DEPTH = 8


def _depth(nodes, P1, P2):
    x, k = nodes // P2, nodes % P2
    return np.minimum(np.minimum(x, P1 - 1 - x), np.minimum(k, P2 - 1 - k))


def _cuts(lo, hi, end):
    """Piece bounds over [0, end) cut every PIECE nodes, only inside [lo, hi]
    (the last piece takes the remainder, up to PIECE + D - 1 nodes)."""
    if end <= 0:
        return []
    b = [0] + [c for c in range(PIECE, end, PIECE) if lo <= c <= hi] + [end]
    return list(zip(b[:-1], b[1:]))


def ring_pieces(P1, P2, D):
    """Output rectangles (r0, r1, c0, c1) covering every node less than D
    from the edge exactly once: top and bottom strips over all columns, then
    left and right strips over the rows between them.  Cuts along a strip
    keep D nodes away from the corners, so the boundary's corner ops (which
    pair nodes two apart along the edge) stay within one piece."""
    out = []
    for c0, c1 in _cuts(D, P2 - D, P2):
        out.append((0, D, c0, c1))
        out.append((P1 - D, P1, c0, c1))
    for r0, r1 in _cuts(0, P1, P1 - 2 * D):
        out.append((D + r0, D + r1, 0, D))
        out.append((D + r0, D + r1, P2 - D, P2))
    return out


def build_ring_plan(ops, P1, P2, ghost):
    """Host arrays of the ring plan for the chunk-ordered op table ``ops``
    (BC_OP_DTYPE, scalar field of padded extents P1 x P2).  Returns a dict
    with depth, nops (the table's op count, which the kernels' entry points
    check against the table they are given), pieces (RING_PIECE_DTYPE), w/t/z (RING_ENTRY_DTYPE), cdst
    (int32), csign (float32)."""
    P1, P2, ghost = int(P1), int(P2), int(ghost)
    if P1 * P2 > 0x3FFFFFFF:
        raise ValueError("field too large for 32-bit byte offsets")
    ops = np.asarray(ops)
    dst = ops["dst"].astype(np.int64) if len(ops) else np.zeros(0, np.int64)
    if len(np.unique(dst)) != len(dst):
        raise ValueError("boundary table writes a scalar twice")
    T = transpose_table(ops) if len(ops) else dict(
        zero_idx=np.zeros(0, np.int64), src_idx=np.zeros(0, np.int64),
        flags=np.zeros(0, np.int32), off=np.zeros(1, np.int64),
        cdst=np.zeros(0, np.int64), csign=np.zeros(0, np.float32))
    reads = ops[ops["kind"] != N.BC_SET] if len(ops) else ops
    touched = np.concatenate([dst, reads["src"].astype(np.int64) if len(reads) else
                              np.zeros(0, np.int64)])
    if touched.size and (touched.min() < 0 or touched.max() >= P1 * P2):
        raise ValueError("boundary table leaves the field")
    dmax = int(_depth(touched, P1, P2).max()) if touched.size else 0
    D = max(4, ghost + 2, dmax + 2)
    if D > MAX_DEPTH:
        raise ValueError(f"boundary reaches {dmax} nodes into the field (at most "
                         f"{MAX_DEPTH - 2} supported)")
    if P1 < 2 * D or P2 < 2 * D:
        raise ValueError("field too small for the boundary ring")

    slot_of = {int(d): m for m, d in enumerate(dst)}
    src_idx, off, cdst = T["src_idx"], T["off"], T["cdst"]

    def rect_mask(nodes, r0, r1, c0, c1):
        x, k = nodes // P2, nodes % P2
        return (x >= r0) & (x < r1) & (k >= c0) & (k < c1)

    pieces, W, Tt, Z = [], [], [], []
    zero_idx = T["zero_idx"]
    dst_nodes = np.array(sorted(slot_of), dtype=np.int64)
    for (r0, r1, c0, c1) in ring_pieces(P1, P2, D):
        g0, g1 = max(r0 - 1, 0), min(r1 + 1, P1)
        h0, h1 = max(c0 - 1, 0), min(c1 + 1, P2)
        if (min(r1 + 2, P1) - max(r0 - 2, 0)) * (min(c1 + 2, P2) - max(c0 - 2, 0)) > RING_CAP:
            raise ValueError("ring piece exceeds the kernel's LDS staging")
        rec = np.zeros(1, N.RING_PIECE_DTYPE)[0]
        rec["r0"], rec["r1"], rec["c0"], rec["c1"] = r0, r1, c0, c1
        rec["w_beg"] = len(W)
        m = rect_mask(dst_nodes, g0, g1, h0, h1)
        inO = rect_mask(dst_nodes, r0, r1, c0, c1)
        for nd, o in zip(dst_nodes[m], inO[m]):
            W.append((int(nd), slot_of[int(nd)], 0, 2 if o else 0))
        rec["w_end"] = len(W)
        rec["t_beg"] = len(Tt)
        m = rect_mask(src_idx, g0, g1, h0, h1)
        inO = rect_mask(src_idx, r0, r1, c0, c1)
        for i in np.flatnonzero(m):
            q0, q1 = int(off[i]), int(off[i + 1])
            cd = cdst[q0:q1]
            if not rect_mask(cd, g0, g1, h0, h1).all():
                raise ValueError("a boundary transpose crosses ring pieces")
            if inO[i] and not rect_mask(cd, r0, r1, c0, c1).all():
                raise ValueError("a boundary transpose crosses ring pieces")
            Tt.append((int(src_idx[i]), q0, q1, int(T["flags"][i]) | (2 if inO[i] else 0)))
        rec["t_end"] = len(Tt)
        if rec["t_end"] - rec["t_beg"] > RING_CAP:
            raise ValueError("too many boundary transpose sources in a ring piece")
        rec["z_beg"] = len(Z)
        m = rect_mask(zero_idx, g0, g1, h0, h1)
        inO = rect_mask(zero_idx, r0, r1, c0, c1)
        for nd, o in zip(zero_idx[m], inO[m]):
            Z.append((int(nd), 0, 0, 2 if o else 0))
        rec["z_end"] = len(Z)
        pieces.append(rec)

    ent = lambda L: np.array(L if L else [(0, 0, 0, 0)], dtype=N.RING_ENTRY_DTYPE)
    return dict(depth=D, nops=len(ops), pieces=np.array(pieces, dtype=N.RING_PIECE_DTYPE), w=ent(W),
                t=ent(Tt), z=ent(Z),
                cdst=np.ascontiguousarray(cdst if len(cdst) else np.zeros(1), dtype=np.int32),
                csign=np.ascontiguousarray(T["csign"] if len(T["csign"]) else np.zeros(1),
                                           dtype=np.float32))


class DeviceRingPlan:
    """The plan's arrays on the device and the ``pde_ring_plan`` record that
    points at them (pass ``ptr`` to the kernel's entry point)."""

    def __init__(self, plan, device):
        import torch
        up = lambda a: torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).copy()).to(device)
        self.depth = plan["depth"]
        self.npieces = len(plan["pieces"])
        self.nops = int(plan["nops"])
        self._arrs = {k: up(plan[k]) for k in ("pieces", "w", "t", "z", "cdst", "csign")}
        rec = np.zeros(1, N.RING_PLAN_DTYPE)
        rec[0] = (self.depth, self.npieces, self.nops,
                  *(self._arrs[k].data_ptr() for k in ("pieces", "w", "t", "z", "cdst", "csign")))
        self.rec = rec

    @property
    def ptr(self):
        return self.rec.ctypes.data


def check_step_sources(plan, ops, P1, P2):
    """For the one-launch heat step (``pde_heat_step_ring_f32``): every
    boundary op a piece evaluates (its written node in the piece grown by one)
    must read a node inside the piece grown by two, and not a written node
    that another piece stores back concurrently.  Raises ``ValueError``."""
    P1, P2 = int(P1), int(P2)
    pieces, W = plan["pieces"], plan["w"]
    owner = {}
    for pid, p in enumerate(pieces):
        for e in W[p["w_beg"]:p["w_end"]]:
            if e["flags"] & 2:
                owner[int(e["node"])] = pid
    for pid, p in enumerate(pieces):
        r0, r1 = max(int(p["r0"]) - 2, 0), min(int(p["r1"]) + 2, P1)
        c0, c1 = max(int(p["c0"]) - 2, 0), min(int(p["c1"]) + 2, P2)
        for e in W[p["w_beg"]:p["w_end"]]:
            op = ops[int(e["a"])]
            if int(op["kind"]) == N.BC_SET:
                continue
            src = int(op["src"])
            x, k = divmod(src, P2)
            if not (r0 <= x < r1 and c0 <= k < c1):
                raise ValueError("a boundary op reads outside its ring piece")
            if owner.get(src, pid) != pid:
                raise ValueError("a boundary op reads a scalar another ring piece writes")
