"""Tile plans for several heat steps per launch (``pde_heat_multistep_f32``).

Each grid (one field, or each instance of a batch) is cut into tiles whose
regions fit the kernel's LDS array (at most ``HEAT_TILE_ROWS`` x
``HEAT_TILE_PITCH`` cells).  Along an axis longer than that, the outputs are
split into pieces and each region is its output grown by ``halo`` cells on the
sides where it was cut from the grid (never past the grid's edge): after
``halo`` sub-steps the cells within ``halo`` of a cut side are stale, the
output is not.  A grid that fits one region is one tile with no cut at all,
stepped whole in LDS.

Each tile gets the boundary ops whose written scalar lies in its region, with
region-local cell indices.  Every op's source must lie in the region and stay
valid as long as its destination is needed: with ``m(c)`` the distance of a
cell from the nearest cut side (cells on a cut side have m = 0; infinite
without cuts), ``m(src) >= min(m(dst), halo - 1)``.  The boundary tables of
GridBoundary (ghost from the adjacent node across the face, Dirichlet
nodes) always pass; a table that breaks it, or has values given by a Python
function (evaluated once per step on the host side), raises ``PlanError`` and
callers keep one step per launch.
"""
from __future__ import annotations

import math

import numpy as np

from . import _native as N

PITCH = N.HEAT_TILE_PITCH
ROWS = N.HEAT_TILE_ROWS
MAXOPS = N.HEAT_TILE_MAXOPS


class PlanError(ValueError):
    pass


def split_axis(P: int, cap: int, halo: int):
    """[(o0, o1, r0, r1)]: output pieces covering [0, P) and their regions."""
    P, cap, halo = int(P), int(cap), int(halo)
    if P <= cap:
        return [(0, P, 0, P)]
    if cap - 2 * halo < 1:
        raise PlanError(f"a halo of {halo} leaves no output in a {cap}-cell region")
    for m in range(2, P + 1):
        L = math.ceil(P / m)
        pieces = []
        for j in range(m):
            o0, o1 = j * L, min(P, (j + 1) * L)
            if o0 >= o1:
                break
            pieces.append((o0, o1, max(0, o0 - halo), min(P, o1 + halo)))
        if all(r1 - r0 <= cap for _, _, r0, r1 in pieces):
            return pieces
    raise PlanError("cannot split the axis")


def plan_grid(P1: int, P2: int, ops: np.ndarray, halo: int, grid: int = 0, op_base: int = 0,
              rows_cap: int = ROWS, cols_cap: int = PITCH):
    """Tiles (HEAT_TILE_DTYPE) and tile ops (HEAT_TILE_OP_DTYPE, op indices
    op_base + index into ``ops``) of one scalar 2D grid of padded extents
    P1 x P2 (column index fastest) whose boundary is the resolved op table
    ``ops`` (BC_OP_DTYPE, flat indices into the grid).  Regions are at most
    rows_cap x cols_cap cells (smaller regions give a small grid more
    workgroups, at the price of more halo cells)."""
    P1, P2, halo = int(P1), int(P2), int(halo)
    if not (1 <= rows_cap <= ROWS and 1 <= cols_cap <= PITCH):
        raise PlanError(f"region caps must be at most {ROWS} x {PITCH}")
    if halo < 1:
        raise PlanError("halo must be at least 1")
    ops = np.asarray(ops)
    if len(ops) and np.any(ops["vkind"] == N.BCV_ARRAY):
        raise PlanError("boundary values given by Python functions are evaluated once per "
                        "step; several steps per launch cannot take them")
    dst = ops["dst"].astype(np.int64) if len(ops) else np.zeros(0, np.int64)
    src = ops["src"].astype(np.int64) if len(ops) else np.zeros(0, np.int64)
    is_set = ops["kind"] == N.BC_SET if len(ops) else np.zeros(0, bool)
    touched = np.concatenate([dst, src[~is_set]])
    if touched.size and (touched.min() < 0 or touched.max() >= P1 * P2):
        raise PlanError("boundary table leaves the grid")
    dx_, dy_ = np.divmod(dst, P2)
    sx_, sy_ = np.divmod(src, P2)
    tiles, tops = [], []
    for (or0, or1, r0, r1) in split_axis(P1, rows_cap, halo):
        for (oc0, oc1, c0, c1) in split_axis(P2, cols_cap, halo):
            rh, rw = r1 - r0, c1 - c0
            sel = np.flatnonzero((dx_ >= r0) & (dx_ < r1) & (dy_ >= c0) & (dy_ < c1))
            inf = np.int64(1 << 40)

            def m(x, y):
                out = np.full(x.shape, inf, dtype=np.int64)
                if r0 > 0:
                    out = np.minimum(out, x - r0)
                if r1 < P1:
                    out = np.minimum(out, r1 - 1 - x)
                if c0 > 0:
                    out = np.minimum(out, y - c0)
                if c1 < P2:
                    out = np.minimum(out, c1 - 1 - y)
                return out

            s_set = is_set[sel]
            sx, sy = sx_[sel], sy_[sel]
            rd = ~s_set
            inside = (sx >= r0) & (sx < r1) & (sy >= c0) & (sy < c1)
            if np.any(rd & ~inside):
                raise PlanError("a boundary op reads outside its tile's region")
            md, ms = m(dx_[sel], dy_[sel]), m(sx, sy)
            if np.any(rd & (ms < np.minimum(md, halo - 1))):
                raise PlanError("a boundary op reads a cell that goes stale before its "
                                "destination does")
            if len(sel) > MAXOPS:
                raise PlanError(f"{len(sel)} boundary ops in one tile (at most {MAXOPS})")
            rec = np.zeros(1, N.HEAT_TILE_DTYPE)[0]
            n0 = sum(len(t) for t in tops)
            rec["grid"], rec["r0"], rec["c0"], rec["rh"], rec["rw"] = grid, r0, c0, rh, rw
            rec["o_r0"], rec["o_r1"], rec["o_c0"], rec["o_c1"] = or0, or1, oc0, oc1
            rec["op_beg"], rec["op_end"] = n0, n0 + len(sel)
            t = np.zeros(len(sel), N.HEAT_TILE_OP_DTYPE)
            t["ldst"] = (dx_[sel] - r0) * PITCH + (dy_[sel] - c0)
            t["lsrc"] = np.where(s_set, -1, (sx - r0) * PITCH + (sy - c0))
            t["op"] = op_base + sel
            tiles.append(rec)
            tops.append(t)
    return np.array(tiles, dtype=N.HEAT_TILE_DTYPE), \
        (np.concatenate(tops) if tops else np.zeros(0, N.HEAT_TILE_OP_DTYPE))


def plan_batch(p1s, p2s, op_tables, halo: int, rows_cap: int = ROWS, cols_cap: int = PITCH):
    """Tiles and tile ops of many grids: grid i's tiles carry grid = i, op
    indices into the concatenation of ``op_tables`` (the packed table)."""
    tiles, tops, base, nt = [], [], 0, 0
    for i, (P1, P2, ops) in enumerate(zip(p1s, p2s, op_tables)):
        t, o = plan_grid(P1, P2, ops, halo, grid=i, op_base=base, rows_cap=rows_cap,
                         cols_cap=cols_cap)
        t["op_beg"] += nt
        t["op_end"] += nt
        nt += len(o)
        base += len(ops)
        tiles.append(t)
        tops.append(o)
    return np.concatenate(tiles), (np.concatenate(tops) if nt else
                                   np.zeros(1, N.HEAT_TILE_OP_DTYPE))


def check_coverage(tiles, p1s, p2s) -> None:
    """Every cell of every grid lies in exactly one output rectangle, each
    output inside its region, each region inside its grid (raises PlanError)."""
    for i, (P1, P2) in enumerate(zip(p1s, p2s)):
        cnt = np.zeros((int(P1), int(P2)), np.int32)
        for t in tiles[tiles["grid"] == i]:
            if not (0 <= t["r0"] <= t["o_r0"] < t["o_r1"] <= t["r0"] + t["rh"] <= P1 and
                    0 <= t["c0"] <= t["o_c0"] < t["o_c1"] <= t["c0"] + t["rw"] <= P2):
                raise PlanError("tile output outside its region or region outside the grid")
            if t["rh"] > ROWS or t["rw"] > PITCH:
                raise PlanError("tile region exceeds the LDS array")
            cnt[t["o_r0"]:t["o_r1"], t["o_c0"]:t["o_c1"]] += 1
        if not np.all(cnt == 1):
            raise PlanError(f"grid {i}: outputs do not cover the grid exactly once")
