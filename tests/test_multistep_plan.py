"""Tile plans for several heat steps per launch (CPU only).

A NumPy emulation of the kernel (``pde_heat_multistep_f32``: region loaded
once, per sub-step the tile's boundary ops two-phase then the 5-point step
in the kernel's FP32 order, output rectangle stored) run on the planner's
tiles must equal k sequential steps of the oracle (boundary, Laplacian,
ForwardEuler) bit for bit: for grids that fit one region (no cuts) and grids
cut in both directions, legacy and functional tables with Neumann, ramp and
sine values, ghost widths 1 and 2, shifted groups.  The planner's refusals
(function values, an op reading across a cut) and its coverage check are
tested too."""
import numpy as np
import pytest
import torch

from helpers import apply_ops_host
from oracle import np_oracle as O
from pde_module_amd import _native as N
from pde_module_amd.deprecated import GridBoundary as LegacyGridBoundary
from pde_module_amd.FDM import GridBoundary
from pde_module_amd.FDM.boundary import get_ramp_func, get_sine_func
from pde_module_amd.multistep import PITCH, PlanError, check_coverage, plan_batch, plan_grid

F = np.float32


def _op_value(op, t):
    if op["vkind"] == N.BCV_SINE:
        return op["value"] * F(np.sin(np.float64(op["tmax"] * F(t))))
    if op["tmax"] > 0:
        return min(F(t) / op["tmax"], F(1.0)) * op["value"]
    return op["value"]


def emulate(u, ops, tiles, tops, P1, P2, g, w, alpha, dt, steps, t=0.0):
    """The kernel's arithmetic, tile by tile (u: (P1, P2) float32)."""
    un = np.full_like(u, np.nan)
    wm, wc, wp = (F(x) for x in w)
    alpha, dt = F(alpha), F(dt)
    for T in tiles:
        r0, c0, rh, rw = (int(T[k]) for k in ("r0", "c0", "rh", "rw"))
        A = u[r0:r0 + rh, c0:c0 + rw].copy()
        mine = tops[T["op_beg"]:T["op_end"]]
        vals = np.array([_op_value(ops[m["op"]], t) for m in mine], dtype=F)
        x = np.arange(r0, r0 + rh)[:, None]
        k = np.arange(c0, c0 + rw)[None, :]
        lr, lc = np.arange(rh)[:, None], np.arange(rw)[None, :]
        step = ((x >= g) & (x < P1 - g) & (k >= g) & (k < P2 - g) & (lr > 0) & (lr < rh - 1)
                & (lc > 0) & (lc < rw - 1))
        for _ in range(steps):
            new = []
            for m, v in zip(mine, vals):
                kind = ops[m["op"]]["kind"]
                if kind == N.BC_SET:
                    new.append(v)
                    continue
                src = A[m["lsrc"] // PITCH, m["lsrc"] % PITCH]
                if kind == N.BC_COPY:
                    new.append(src)
                elif kind == N.BC_GHOST_D:
                    new.append(F(2.0) * v - src)
                else:
                    new.append(ops[m["op"]]["coef"] * v + src)
            for m, v in zip(mine, new):
                A[m["ldst"] // PITCH, m["ldst"] % PITCH] = v
            up = np.zeros_like(A)
            up[1:] = A[:-1]
            dn = np.zeros_like(A)
            dn[:-1] = A[1:]
            zl = np.zeros_like(A)
            zl[:, 1:] = A[:, :-1]
            zr = np.zeros_like(A)
            zr[:, :-1] = A[:, 1:]
            acc = F(0.0) + wm * up
            acc = acc + wc * A
            acc = acc + wp * dn
            acc = acc + wm * zl
            acc = acc + wc * A
            acc = acc + wp * zr
            A = np.where(step, A + dt * (alpha * acc), A)
        o = (slice(T["o_r0"] - r0, T["o_r1"] - r0), slice(T["o_c0"] - c0, T["o_c1"] - c0))
        un[T["o_r0"]:T["o_r1"], T["o_c0"]:T["o_c1"]] = A[o]
    return un


def sequential(u, ops, g, w, alpha, dt, steps, t=0.0):
    """steps x (boundary ops in place, then the heat step), the oracle's
    Laplacian and Euler update on interior nodes, u2 elsewhere."""
    P1, P2 = u.shape
    u = u.reshape(P1, P2, 1, 1)
    m = O.interior_mask((P1, P2, 1), g)[..., None]
    for _ in range(steps):
        u2 = apply_ops_host(u, ops, t)
        lap = O.laplacian(u2, w, F(alpha), g)
        u = np.where(m, O.forward_euler(u2, np.where(m, lap, F(0)), F(dt)), u2)
    return u.reshape(P1, P2)


def _legacy_mixed(u, dx, g):
    bc = LegacyGridBoundary(u, dx, g)
    bc.vonNeumann_BC("-X", 0.5)
    bc.dirichlet_BC("+Y", 0.75)
    return bc


def _legacy_shift(u, dx, g):
    bc = LegacyGridBoundary(u, dx, g)
    bc.dirichlet_BC("ALL", 0.25)
    bc.shift_group("-X", 3, 0)
    return bc


def _functional(u, dx, g):
    bc = GridBoundary(u, dx, g)
    bc.dirichlet_BC("ALL", get_ramp_func(2.0, 1.0))
    bc.vonNeumann_BC("+Y", 0.25)
    bc.dirichlet_BC("-Y", get_sine_func(1.5, 7.0))
    return bc


CASES = [((40, 37), 1, _legacy_mixed, 8), ((128, 128), 1, _functional, 4),
         ((300, 140), 1, _legacy_mixed, 8), ((133, 270), 2, _functional, 5),
         ((260, 261), 1, _legacy_shift, 16), ((129, 40), 1, _legacy_shift, 3)]


@pytest.mark.parametrize("caps", [(128, 128), (64, 48)])
@pytest.mark.parametrize("shape,g,make,k", CASES)
def test_emulated_tiles_equal_sequential_steps(shape, g, make, k, caps):
    P1, P2 = shape
    rng = np.random.default_rng(P1 * P2)
    u = rng.uniform(-1, 1, (P1, P2)).astype(F)
    dx, alpha = 0.01, 0.05
    dt = 0.2 * dx * dx / alpha
    bc = make(torch.from_numpy(u.reshape(P1, P2, 1, 1)), dx, g)
    ops = bc.resolved_ops()
    if 2 * k >= min(caps):
        with pytest.raises(PlanError):
            plan_grid(P1, P2, ops, k, rows_cap=caps[0], cols_cap=caps[1])
        return
    tiles, tops = plan_grid(P1, P2, ops, k, rows_cap=caps[0], cols_cap=caps[1])
    check_coverage(tiles, [P1], [P2])
    assert np.all(tiles["rh"] <= caps[0]) and np.all(tiles["rw"] <= caps[1])
    if P1 <= caps[0] and P2 <= caps[1]:
        assert len(tiles) == 1 and tiles["rh"][0] == P1 and tiles["rw"][0] == P2
    w = O.laplacian_weights(F(dx))
    for steps in (k, max(1, k // 2)):      # a plan for k serves any count up to k
        got = emulate(u, ops, tiles, tops, P1, P2, g, w, alpha, dt, steps, t=0.3)
        want = sequential(u, ops, g, w, alpha, dt, steps, t=0.3)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), steps


def test_batch_plan_indexes_the_packed_table():
    shapes = [(40, 37), (300, 140), (9, 9)]
    tabs = [_legacy_mixed(torch.zeros(s + (1, 1)), 0.01, 1).resolved_ops() for s in shapes]
    tiles, tops = plan_batch([s[0] for s in shapes], [s[1] for s in shapes], tabs, 8)
    check_coverage(tiles, [s[0] for s in shapes], [s[1] for s in shapes])
    base = np.cumsum([0] + [len(t) for t in tabs])
    for T in tiles:
        mine = tops[T["op_beg"]:T["op_end"]]
        gi = T["grid"]
        assert np.all((mine["op"] >= base[gi]) & (mine["op"] < base[gi + 1]))


def test_planner_refusals():
    u = torch.zeros((300, 140, 1, 1))
    bc = GridBoundary(u, 0.01, 1)
    bc.dirichlet_BC("ALL", lambda *a: 1.0)
    with pytest.raises(PlanError):
        plan_grid(300, 140, bc.resolved_ops(), 8)
    ops = _legacy_mixed(u, 0.01, 1).resolved_ops().copy()
    # a ghost op of the -Y face (column 0) on row 105, which the regions of
    # the first two row pieces ([0, 108) and [92, 208)) both hold, made to read
    # row 130: outside the first region
    i = int(np.flatnonzero((ops["kind"] != N.BC_SET) & (ops["dst"] == 105 * 140))[0])
    ops["src"][i] = 130 * 140 + 1
    with pytest.raises(PlanError):
        plan_grid(300, 140, ops, 8)
    # inside the region but stale too early: row 104 lies 12 rows inside the
    # second piece's cut at row 92; a source on row 93 goes stale after one
    # sub-step while the destination is needed for 7
    ops2 = _legacy_mixed(u, 0.01, 1).resolved_ops().copy()
    i = int(np.flatnonzero((ops2["kind"] != N.BC_SET) & (ops2["dst"] == 104 * 140))[0])
    ops2["src"][i] = 93 * 140 + 1
    with pytest.raises(PlanError):
        plan_grid(300, 140, ops2, 8)
    tiles, _ = plan_grid(300, 140, _legacy_mixed(u, 0.01, 1).resolved_ops(), 8)
    bad = tiles.copy()
    bad["o_r1"][0] -= 1
    with pytest.raises(PlanError):
        check_coverage(bad, [300], [140])
