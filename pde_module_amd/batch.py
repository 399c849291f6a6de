"""Batched ensembles of independent 2D heat grids of uneven sizes.

The reference steps one grid per module call; an ensemble there is a Python
loop over instances, each step three launches (boundary, Laplacian,
ForwardEuler: ``examples/2D_Heat_Equation.py:58-62``).  ``HeatEnsemble`` packs
every instance's padded field into one buffer and steps all of them with two
launches per step:

1. ``pde_bc_apply_chunked_inplace_f32`` over one packed op table: every
   instance's ops in conflict-closed chunks of at most 256, one workgroup per
   chunk (``pde_bc_apply_batch_inplace_f32``, one workgroup per instance, when
   some table cannot be chunked);
2. ``pde_heat_batch_step_f32`` — one tile list over all instances; narrow
   instances use row segments of 32 or 16 lanes so most lanes stay busy.

The time loop is captured once as a HIP graph of ``graph_chunk`` steps and
replayed.  The tile shape (rows per lane, waves per block) and the launch
mode (one block per tile, or a persistent grid looping over the tiles) are
chosen by ``autotune`` from timings of the actual ensemble.

Every instance ends bit-identical to stepping it alone with ``HeatStep`` (the
same boundary ops and the same per-node arithmetic).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _native as N
from .FDM.module.laplacian import laplacian_weights
from .time_step import HeatStep

ALIGN = 64            # floats: every instance starts on a 256-byte boundary
ROWS_PER_LANE = (1, 2, 4, 8)
WAVES = (1, 2, 4)
MULTISTEP_CANDIDATES = (4, 8, 12, 16)


def _plane(shape3):
    """(p1, p2) of a field with exactly two grid axes of size > 1."""
    ext = [int(s) for s in shape3 if s > 1]
    if len(ext) != 2:
        raise ValueError(f"batched heat grids are 2D; got grid shape {tuple(shape3)}")
    return ext[0], ext[1]


def segment_width(nq: int) -> int:
    """Lanes per row segment for a row of nq 4-column chunks: the width in
    {64, 32, 16} with the fewest idle lanes (ties go to the wider one)."""
    best, best_util = 64, -1.0
    for cw in (64, 32, 16):
        util = nq / (math.ceil(nq / cw) * cw)
        if util > best_util + 1e-12:
            best, best_util = cw, util
    return best


def build_tiles(p1s, p2s, rows_per_lane: int, waves: int, xcd_slots: int = 8) -> np.ndarray:
    """Tile list (int32, 4 per tile: grid, first row, first chunk, cw) over
    every grid.

    Blocks are dispatched round-robin over the chip's XCDs (block b to XCD
    b mod 8), each with its own L2.  A tile re-reads one halo row above and
    below, owned by the tiles above and below it in the same column strip; the
    list is therefore laid out so that each column strip's tiles, top to
    bottom, sit at block indices congruent mod ``xcd_slots``: one L2 serves a
    strip and the halo rows are hits.  Strips are taken largest grid first;
    ``xcd_slots=1`` gives plain row-major order."""
    order = sorted(range(len(p1s)), key=lambda i: -(p1s[i] * p2s[i]))
    strips = []
    for gi in order:
        p1, p2 = int(p1s[gi]), int(p2s[gi])
        nq = (p2 + 3) // 4
        cw = segment_width(nq)
        rows = waves * (64 // cw) * rows_per_lane
        x0 = np.arange(0, p1, rows, dtype=np.int32)
        for q0 in range(0, nq, cw):
            t = np.empty((len(x0), 4), np.int32)
            t[:, 0], t[:, 1], t[:, 2], t[:, 3] = gi, x0, q0, cw
            strips.append(t)
    if not strips:
        return np.zeros((0, 4), np.int32)
    if xcd_slots <= 1:
        return np.concatenate(strips)
    # slot s emits the tiles of its current strip one at a time, in turn with
    # the other slots; a finished strip is replaced by the next one.  Once the
    # strips run out, idle slots emit empty tiles (grid -1, skipped by the
    # kernel) so the alignment holds to the end.
    out = []
    queue = iter(strips)
    slots = [next(queue, None) for _ in range(xcd_slots)]
    pos = [0] * xcd_slots
    empty = np.array([-1, 0, 0, 64], np.int32)
    while True:
        for k in range(xcd_slots):
            while slots[k] is not None and pos[k] >= len(slots[k]):
                slots[k], pos[k] = next(queue, None), 0
        if all(sl is None for sl in slots):
            break
        for k in range(xcd_slots):
            if slots[k] is None:
                out.append(empty)
            else:
                out.append(slots[k][pos[k]])
                pos[k] += 1
    return np.stack(out).astype(np.int32)


@dataclass
class TileConfig:
    rows_per_lane: int = 4
    waves: int = 4
    blocks: int = 0          # 0: one block per tile; else a persistent grid
    xcd_slots: int = 8       # tile order: strips aligned to XCDs (8) or row-major (1)
    # > 1: that many steps per launch on LDS-resident tiles (pde_heat_multistep_f32,
    # the boundary applied inside); the other fields then do not apply
    steps_per_launch: int = 1

    def key(self):
        return (self.rows_per_lane, self.waves, self.blocks, self.xcd_slots,
                self.steps_per_launch)


class HeatEnsemble:
    """``HeatEnsemble(boundaries, dxs, ghost_cells, alphas, dts)``: one
    GridBoundary (functional or legacy) per instance, built for its
    ``(X, Y, 1, 1)`` field; per-instance dx, diffusivity and time step.

    ``load(fields)`` copies the initial fields in; ``run(steps)`` advances
    every instance; ``fields()`` returns views of the current states (same
    shapes as loaded)."""

    def __init__(self, boundaries, dxs, ghost_cells: int, alphas, dts, device=None,
                 graph_chunk: int = 50, tiles: TileConfig | None = None):
        self.n = len(boundaries)
        if not (len(dxs) == len(alphas) == len(dts) == self.n) or self.n == 0:
            raise ValueError("one dx, alpha and dt per boundary")
        self.g = int(ghost_cells)
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        self.shapes = [tuple(int(s) for s in bc.grid_shape) + (1,) for bc in boundaries]
        for bc in boundaries:
            if int(bc.inputs[0]) != 1:
                raise ValueError("batched heat grids are scalar fields")
        planes = [_plane(s[:3]) for s in self.shapes]
        self.p1 = np.array([p[0] for p in planes], np.int64)
        self.p2 = np.array([p[1] for p in planes], np.int64)
        sizes = self.p1 * self.p2
        offs = np.zeros(self.n, np.int64)
        pos = 0
        for i in range(self.n):
            offs[i] = pos
            pos += int(math.ceil(sizes[i] / ALIGN) * ALIGN)
        self.total = pos
        self.offsets = offs
        self.points = int(sum((p1 - 2 * self.g) * (p2 - 2 * self.g)
                              for p1, p2 in zip(self.p1, self.p2)))

        # packed boundary table, every index shifted to the packed buffer, each
        # instance's ops in conflict-closed chunk order
        from .FDM.boundary.tables import chunk_ops
        cap = N.load().pde_batch_bc_capacity()
        tables, begin = [], 0
        self._local_ops = []       # per instance, chunk order, indices into its own field
        chunk_starts, chunked = [np.zeros(1, np.int64)], True
        grid = np.zeros(self.n, N.BATCH_GRID_DTYPE)
        for i, bc in enumerate(boundaries):
            ops = np.array(bc.resolved_ops(), dtype=N.BC_OP_DTYPE, copy=True)
            if np.any(ops["vkind"] == N.BCV_ARRAY):
                raise ValueError(f"instance {i}: boundary values given by Python functions are "
                                 "not supported in a batched ensemble (use constant, ramp or "
                                 "sine values)")
            ch = chunk_ops(ops) if len(ops) else (ops, np.zeros(1, np.int64))
            if ch is None:
                chunked = False
            else:
                ops = np.array(ch[0], copy=True)
                chunk_starts.append(ch[1][1:] + begin)
            if not chunked and len(ops) > cap:
                raise ValueError(f"instance {i} has {len(ops)} boundary ops; at most {cap}")
            rd = ops["src"][ops["kind"] != N.BC_SET]
            idx = np.concatenate([ops["dst"], rd]) if len(ops) else np.zeros(0, np.int64)
            if len(idx) and (idx.min() < 0 or idx.max() >= sizes[i]):
                raise IndexError(f"instance {i}: boundary op index outside its field")
            self._local_ops.append(ops.copy())
            ops["dst"] += offs[i]
            ops["src"] += offs[i]
            tables.append(ops)
            wm, wc, wp = laplacian_weights(dxs[i])
            grid[i] = (offs[i], begin, begin + len(ops), self.p1[i], self.p2[i], self.g, 0,
                       wm, wc, wp, np.float32(alphas[i]), np.float32(dts[i]), 0.0)
            begin += len(ops)
        self.host_grid = grid
        allops = np.concatenate(tables) if begin else np.zeros(1, N.BC_OP_DTYPE)
        self.needs_t = bool(np.any((allops["tmax"] > 0) | (allops["vkind"] != N.BCV_CONST)))
        dev = self.device
        self.grid_dev = torch.from_numpy(grid.view(np.uint8).copy()).to(dev)
        self.ops_dev = torch.from_numpy(allops.view(np.uint8).copy()).to(dev)
        self.chunk_start = None
        if chunked:
            cs = np.concatenate(chunk_starts).astype(np.int64)
            self.chunk_start = torch.from_numpy(cs).to(dev)
            self.nchunks = len(cs) - 1
        self.t = torch.zeros(1, dtype=torch.float32, device=dev)
        self.bufs = [torch.zeros(self.total, dtype=torch.float32, device=dev) for _ in range(2)]
        self.cur = 0
        self.graph_chunk = max(2, int(graph_chunk) // 2 * 2)
        self._graph = None
        self._tile_cache = {}
        self._mplans = {}
        self.config = tiles or TileConfig()

    # ----------------------------------------------------------------- state
    def load(self, fields) -> None:
        if len(fields) != self.n:
            raise ValueError("one field per instance")
        buf = self.bufs[self.cur]
        for i, f in enumerate(fields):
            f = torch.as_tensor(f)
            if tuple(f.shape) != self.shapes[i]:
                raise AssertionError(f"instance {i}: shape {tuple(f.shape)} != {self.shapes[i]}")
            n = int(self.p1[i] * self.p2[i])
            buf[int(self.offsets[i]):int(self.offsets[i]) + n].copy_(
                f.reshape(-1).to(device=buf.device, dtype=torch.float32))

    def fields(self):
        buf = self.bufs[self.cur]
        out = []
        for i in range(self.n):
            n = int(self.p1[i] * self.p2[i])
            o = int(self.offsets[i])
            out.append(buf[o:o + n].view(self.shapes[i]))
        return out

    # ----------------------------------------------------------------- steps
    def _tiles(self, cfg: TileConfig):
        key = (cfg.rows_per_lane, cfg.waves, cfg.xcd_slots)
        if key not in self._tile_cache:
            if cfg.rows_per_lane not in ROWS_PER_LANE or cfg.waves not in WAVES:
                raise ValueError(f"unsupported tile config {cfg}")
            t = build_tiles(self.p1, self.p2, cfg.rows_per_lane, cfg.waves, cfg.xcd_slots)
            self._tile_cache[key] = (torch.from_numpy(t.reshape(-1).copy()).to(self.device),
                                     len(t))
        return self._tile_cache[key]

    def _mstep_plan(self, halo: int):
        """Device tiles and tile ops of the multi-step launch with this halo
        (cached); raises multistep.PlanError when some instance's boundary
        cannot take it."""
        from .multistep import check_coverage, plan_batch
        if halo not in self._mplans:
            tiles, tops = plan_batch(self.p1, self.p2, self._local_ops, halo)
            check_coverage(tiles, self.p1, self.p2)
            dev = lambda a: torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).copy()).to(
                self.device)
            self._mplans[halo] = (dev(tiles), len(tiles), dev(tops))
        return self._mplans[halo]

    def multistep_ok(self, steps_per_launch: int) -> bool:
        """Whether every instance's boundary allows this many steps per launch."""
        from .multistep import PlanError
        try:
            self._mstep_plan(int(steps_per_launch))
            return True
        except PlanError:
            return False

    def _launch_multistep(self, src, dst, k: int, halo: int) -> None:
        tiles, nt, tops = self._mstep_plan(halo)
        N.call("pde_heat_multistep_f32", src.data_ptr(), dst.data_ptr(), self.grid_dev.data_ptr(),
               self.n, tiles.data_ptr(), nt, tops.data_ptr(), self.ops_dev.data_ptr(),
               self.t.data_ptr() if self.needs_t else None, int(k), N.stream_of(src))

    def _step(self, src: torch.Tensor, dst: torch.Tensor, cfg: TileConfig) -> None:
        st = N.stream_of(src)
        tp = self.t.data_ptr() if self.needs_t else None
        if self.chunk_start is not None:
            N.call("pde_bc_apply_chunked_inplace_f32", src.data_ptr(), self.ops_dev.data_ptr(),
                   self.chunk_start.data_ptr(), self.nchunks, tp, st)
        else:
            N.call("pde_bc_apply_batch_inplace_f32", src.data_ptr(), self.grid_dev.data_ptr(),
                   self.n, self.ops_dev.data_ptr(), tp, st)
        tiles, nt = self._tiles(cfg)
        N.call("pde_heat_batch_step_f32", src.data_ptr(), dst.data_ptr(),
               self.grid_dev.data_ptr(), self.n, tiles.data_ptr(), nt, cfg.rows_per_lane,
               cfg.waves, cfg.blocks, st)

    def _steps(self, k: int, cfg: TileConfig) -> None:
        """k steps from bufs[cur], one launch (pair of launches) per step or
        steps_per_launch per launch; cur follows the launches."""
        spl = int(cfg.steps_per_launch)
        if spl > 1:
            while k > 0:
                n = min(spl, k)
                self._launch_multistep(self.bufs[self.cur], self.bufs[1 - self.cur], n, spl)
                self.cur = 1 - self.cur
                k -= n
            return
        for _ in range(k):
            self._step(self.bufs[self.cur], self.bufs[1 - self.cur], cfg)
            self.cur = 1 - self.cur

    def _chunk(self, cfg: TileConfig) -> int:
        """Steps per captured graph: an even number of launches, so a replay
        leaves the state where it started."""
        spl = max(1, int(cfg.steps_per_launch))
        per = 2 * spl
        return max(per, (self.graph_chunk + per - 1) // per * per)

    def set_time(self, t: float) -> None:
        """Time the ramp boundary values read (a device scalar, so graphs see it)."""
        self.t.fill_(float(np.float32(t)))

    def run(self, steps: int, use_graph: bool = True) -> None:
        """Advance every instance by ``steps`` steps."""
        steps = int(steps)
        cfg = self.config
        if not use_graph:
            self._steps(steps, cfg)
            return
        k = self._chunk(cfg)
        full = steps // k
        if full:
            if self._graph is None or self._graph[0] != (cfg.key(), self.cur):
                if cfg.steps_per_launch > 1:
                    self._mstep_plan(int(cfg.steps_per_launch))   # upload before capture
                else:
                    self._tiles(cfg)              # upload before capture
                torch.cuda.synchronize(self.device)
                g = torch.cuda.CUDAGraph()
                start = self.cur
                with torch.cuda.graph(g):
                    self._steps(k, cfg)
                self.cur = start              # capture recorded, did not run
                self._graph = ((cfg.key(), start), g)
            for _ in range(full):
                self._graph[1].replay()
        self._steps(steps - full * k, cfg)

    # -------------------------------------------------------------- autotune
    def autotune(self, candidates=None, steps: int = 8, reps: int = 3) -> TileConfig:
        """Times each candidate tile config on this ensemble (state restored
        afterwards) and keeps the fastest; returns it."""
        if candidates is None:
            sms = torch.cuda.get_device_properties(self.device).multi_processor_count
            candidates = [TileConfig(r, w, b, x) for r in ROWS_PER_LANE for w in WAVES
                          for b in (0, 8 * sms) for x in (8, 1)]
            candidates += [TileConfig(steps_per_launch=s) for s in MULTISTEP_CANDIDATES
                           if self.multistep_ok(s)]
        saved = [b.clone() for b in self.bufs]
        cur = self.cur
        best, best_t = None, float("inf")
        timings = {}
        for cfg in candidates:
            if cfg.steps_per_launch > 1:
                self._mstep_plan(int(cfg.steps_per_launch))
            else:
                self._tiles(cfg)
            n = max(steps, 2 * int(cfg.steps_per_launch))
            self._steps(2, cfg)
            times = []
            for _ in range(reps):
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                self._steps(n, cfg)
                e1.record()
                e1.synchronize()
                times.append(e0.elapsed_time(e1) / n)
            timings[cfg.key()] = min(times)
            if min(times) < best_t:
                best, best_t = cfg, min(times)
        for b, s in zip(self.bufs, saved):
            b.copy_(s)
        self.cur = cur
        self.config = best
        self._graph = None
        self.timings = timings
        return best


class PerInstanceHeat:
    """The unbatched baseline: one ``HeatStep`` per instance, stepped in a
    Python loop (two launches per instance and step)."""

    def __init__(self, boundaries, dxs, ghost_cells, alphas, dts):
        self.steps = [HeatStep(bc, dx, ghost_cells, a, dt)
                      for bc, dx, a, dt in zip(boundaries, dxs, alphas, dts)]
        self.state = None

    def load(self, fields) -> None:
        self.state = [s._own(torch.as_tensor(f)) for s, f in zip(self.steps, fields)]

    def run(self, steps: int) -> None:
        for i, st in enumerate(self.steps):
            u = self.state[i]
            for _ in range(int(steps)):
                u = st(u)
            self.state[i] = u

    def fields(self):
        return list(self.state)
