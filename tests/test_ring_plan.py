"""Host checks of the boundary ring plan behind the time-reversed wave
gradient and the one-launch heat step (``pde_module_amd.ring_plan``): the
pieces tile exactly the nodes closer than ``depth`` to the edge, every entry
the kernel applies lies where the piece can reach it, and tables the kernel
cannot handle are refused."""
import numpy as np
import pytest
import torch

from pde_module_amd import _native as N
from pde_module_amd.deprecated import GridBoundary as LegacyGridBoundary
from pde_module_amd.FDM import GridBoundary
from pde_module_amd.FDM.boundary.tables import chunk_ops, transpose_table
from pde_module_amd.ring_plan import RING_CAP, build_ring_plan, check_step_sources, ring_pieces


def _ops(shape, make):
    u = torch.zeros(shape + (1,))
    bc = make(u)
    ch = chunk_ops(bc.resolved_ops())
    return ch[0]


LEGACY = lambda u: LegacyGridBoundary(u, 0.01, 1)


def _mixed(u):
    bc = LegacyGridBoundary(u, 0.01, 1)
    bc.vonNeumann_BC("-X", 0.5)
    bc.vonNeumann_BC("+Y", 0.0)
    return bc


def _functional(u):
    bc = GridBoundary(u, 0.01, 1)
    bc.dirichlet_BC("ALL", 0.0)
    bc.vonNeumann_BC("+Y", 0.25)
    return bc


@pytest.mark.parametrize("shape", [(8, 8, 1), (9, 40, 1), (34, 34, 1), (61, 525, 1),
                                   (300, 77, 1), (2050, 2050, 1)])
@pytest.mark.parametrize("make", [LEGACY, _mixed, _functional])
def test_plan_covers_ring_and_places_entries(shape, make):
    P1, P2 = shape[:2]
    ops = _ops(shape, make)
    plan = build_ring_plan(ops, P1, P2, 1)
    D = plan["depth"]
    assert 4 <= D <= 8
    x, k = np.indices((P1, P2))
    depth = np.minimum(np.minimum(x, P1 - 1 - x), np.minimum(k, P2 - 1 - k))
    cover = np.zeros((P1, P2), int)
    for p in plan["pieces"]:
        assert p["r0"] < p["r1"] and p["c0"] < p["c1"]
        cover[p["r0"]:p["r1"], p["c0"]:p["c1"]] += 1
        grown = (min(p["r1"] + 2, P1) - max(p["r0"] - 2, 0)) * \
                (min(p["c1"] + 2, P2) - max(p["c0"] - 2, 0))
        assert grown <= RING_CAP
    assert np.array_equal(cover, (depth < D).astype(int))
    # every node the boundary or its transpose touches is at most D - 2 deep
    T = transpose_table(ops)
    touched = np.concatenate([ops["dst"], ops["src"][ops["kind"] != N.BC_SET], T["cdst"]])
    assert depth.ravel()[touched].max() <= D - 2
    # entries lie in the piece grown by one; bit 1 marks the piece itself, and
    # transposes of those stay inside the piece
    inr = lambda n, r0, r1, c0, c1: (r0 <= n // P2 < r1) and (c0 <= n % P2 < c1)
    slots = {int(d): m for m, d in enumerate(ops["dst"])}
    seen_w = set()
    for p in plan["pieces"]:
        g = (max(p["r0"] - 1, 0), min(p["r1"] + 1, P1), max(p["c0"] - 1, 0), min(p["c1"] + 1, P2))
        o = (p["r0"], p["r1"], p["c0"], p["c1"])
        for e in plan["w"][p["w_beg"]:p["w_end"]]:
            assert inr(int(e["node"]), *g) and slots[int(e["node"])] == e["a"]
            assert bool(e["flags"] & 2) == inr(int(e["node"]), *o)
            if e["flags"] & 2:
                seen_w.add(int(e["node"]))
        for e in plan["t"][p["t_beg"]:p["t_end"]]:
            cd = plan["cdst"][e["a"]:e["b"]]
            box = o if e["flags"] & 2 else g
            assert all(inr(int(c), *box) for c in cd)
        for e in plan["z"][p["z_beg"]:p["z_end"]]:
            assert inr(int(e["node"]), *g)
    assert seen_w == set(slots)   # every written scalar restored by exactly its owner


def test_plan_rejects_what_the_kernel_cannot_run():
    ops = _ops((34, 34, 1), LEGACY)
    with pytest.raises(ValueError):
        build_ring_plan(ops, 6, 6, 1)            # too small for the ring
    deep = ops.copy()
    deep["src"][deep["kind"] != N.BC_SET] = 17 * 34 + 17   # reads the centre
    with pytest.raises(ValueError):
        build_ring_plan(deep, 34, 34, 1)
    dup = np.concatenate([ops, ops[:1]])
    with pytest.raises(ValueError):
        build_ring_plan(dup, 34, 34, 1)
    # a ghost read from a node two rows in but 60 columns away: shallow, yet
    # its transpose crosses ring pieces
    ops2 = _ops((40, 134, 1), LEGACY)
    far = ops2.copy()
    top = np.flatnonzero((far["kind"] == N.BC_GHOST_D) & (far["dst"] // 134 == 0)
                         & (far["dst"] % 134 == 10))
    assert len(top) == 1
    far["src"][top[0]] = 2 * 134 + 70
    build_ring_plan(ops2, 40, 134, 1)
    with pytest.raises(ValueError):
        build_ring_plan(far, 40, 134, 1)


def test_pieces_keep_corner_pairs_together():
    """Cuts along a strip stay D nodes from the corners, so a last piece is
    never narrower than the corner ops' reach."""
    for P2 in range(8, 140):
        pcs = ring_pieces(40, P2, 4)
        top = [p for p in pcs if p[0] == 0]
        assert all(c0 == 0 or (4 <= c0 <= P2 - 4) for (_, _, c0, _) in top)


@pytest.mark.parametrize("shape", [(9, 40, 1), (34, 34, 1), (61, 525, 1), (300, 77, 1)])
@pytest.mark.parametrize("make", [LEGACY, _mixed, _functional])
def test_step_sources_of_standard_tables_fit_the_one_launch_heat_step(shape, make):
    ops = _ops(shape, make)
    plan = build_ring_plan(ops, shape[0], shape[1], 1)
    check_step_sources(plan, ops, shape[0], shape[1])


def test_step_sources_refuse_a_read_of_a_scalar_another_piece_writes():
    """A ghost on a piece cut reading the next ghost along: the ring plan takes
    it (its transpose stays in the owning piece), but the piece on the other
    side of the cut also evaluates that op and would stage the source while its
    owner writes the boundary value back, so the one-launch heat step refuses
    the table (``HeatStep`` then takes the two-launch path)."""
    P1, P2 = 40, 134
    ops = _ops((P1, P2, 1), LEGACY)
    plan = build_ring_plan(ops, P1, P2, 1)
    check_step_sources(plan, ops, P1, P2)
    cut = min(int(p["c0"]) for p in plan["pieces"] if p["r0"] == 0 and p["c0"] > 0)
    bad = ops.copy()
    i = np.flatnonzero((bad["kind"] != N.BC_SET) & (bad["dst"] == cut))
    assert len(i) == 1
    bad["src"][i[0]] = cut + 1
    plan2 = build_ring_plan(bad, P1, P2, 1)
    with pytest.raises(ValueError):
        check_step_sources(plan2, bad, P1, P2)
