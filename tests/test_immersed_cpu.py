"""Immersed boundary (staircase): the oracle pinned by staircase identities
(Dirichlet values, the Neumann mirror and gradient, the mean of two fluid
neighbours, interior fill) and the module's host-side geometry against the
oracle's (CPU only)."""
import numpy as np
import pytest
import torch

from oracle import np_oracle as O
from pde_module_amd.FDM import ImmersedBoundary


def _block(shape=(12, 10, 1)):
    bm = np.zeros(shape, np.int8)
    bm[4:8, 3:7, 0] = 1          # 4 x 4 solid block
    return bm


def test_oracle_staircase_identities():
    bm = _block()
    rng = np.random.default_rng(0)
    u = rng.uniform(-1, 1, bm.shape + (2,)).astype(np.float64)
    nb = 12                       # boundary nodes of a 4 x 4 block
    btype = np.ones((nb, 2), np.int8)
    bval = np.zeros((nb, 2))
    bval[:, 0] = 3.0
    btype[:, 1] = 2               # Neumann on component 1, value 0.5
    bval[:, 1] = 0.5
    dx = 0.1
    out = O.immersed_boundary(u, bm, btype, bval, dx, fill_value=-7.0)
    # Dirichlet component everywhere on the boundary, interior solids filled
    assert np.all(out[4:8, 3:7, 0, 0][[0, 0, 3, 3], [0, 3, 0, 3]] == 3.0)
    assert np.all(out[5:7, 4:6, 0, :] == -7.0)
    # an edge node (not a corner): one fluid neighbour -> (u_f - u_s)/(s dx) = v
    s_node, f_node = (4, 5, 0), (3, 5, 0)     # fluid on the -X side (s = -1)
    assert np.isclose((u[f_node + (1,)] - out[s_node + (1,)]) / (-dx), 0.5)
    # a corner: mean of its two fluid neighbours' mirrored values
    c_node = (4, 3, 0)
    want = ((u[3, 3, 0, 1] + 0.5 * dx) + (u[4, 2, 0, 1] + 0.5 * dx)) / 2
    assert np.isclose(out[c_node + (1,)], want)
    # fluid untouched
    m = bm == 0
    assert np.array_equal(out[m], u[m])


def test_module_geometry_matches_oracle_and_tables():
    bm = _block()
    field = torch.zeros(bm.shape + (2,))
    ib = ImmersedBoundary(field, 0.1, 1)
    ib.from_bool_func(lambda x, y, z: (x >= 0.35) & (x <= 0.75) & (y >= 0.25) & (y <= 0.65),
                      [np.arange(12)[:, None, None] * 0.1 * np.ones(bm.shape),
                       np.arange(10)[None, :, None] * 0.1 * np.ones(bm.shape),
                       np.zeros(bm.shape)])
    assert np.array_equal(ib.bitmask, bm)
    ib.finalize()
    assert len(ib.solid_boundary) == 12 and len(ib.interior_solids) == 4
    assert ib.fluid_neighbors.shape == (12, 2, 2)
    assert len(ib.fluid_boundary) == 16                 # the ring around the block
    ib.dirichlet_BC("ALL", 1.5, 0)
    ib.vonNeumann_BC("ALL", 0.25, 1)
    ops, fill = ib.resolved_ops()
    assert len(ops) == 24 and len(fill) == 4
    assert set(ops["kind"][ops["comp"] == 0]) == {1} and set(ops["kind"][ops["comp"] == 1]) == {2}
    # the corner node (4, 3) has fluid on -X (axis 0, bit 0) and -Y (axis 1, bit 2)
    k = int(np.nonzero((ops["node"] == 4 * 10 + 3) & (ops["comp"] == 1))[0][0])
    assert ops["mask"][k] == 0b0101


def test_outer_layer_neighbours_use_flat_index_arithmetic():
    """A solid node in ghost row 0 has its -Y neighbour at flat index - stride,
    the previous row's last node; a solid in the first plane's first row
    has none (flat index -stride < 0) and raises."""
    field = torch.zeros((6, 5, 1, 1))
    ib = ImmersedBoundary(field, 0.1, 1)
    ib.bitmask[2, 0, 0] = 1
    ib.bitmask[1, 4, 0] = 1           # the node just before (2, 0) in flat order
    ib.finalize()
    k = [tuple(p) for p in ib.solid_boundary].index((2, 0, 0))
    assert ib.fluid_neighbors[k, 1, 0] == 1          # -Y neighbour is the solid (1, 4)
    ib2 = ImmersedBoundary(field, 0.1, 1)
    ib2.bitmask[0, 0, 0] = 1
    with pytest.raises(ValueError):
        ib2.finalize()


def test_oracle_rejects_what_the_module_rejects():
    """A solid node on the first row of the first plane has its -X and -Y
    neighbours before the array's start: the module's finalize and the oracle
    both raise instead of wrapping the index; a solid in ghost row 0 further in
    is accepted by both and treated alike."""
    u = np.zeros((6, 5, 1, 1))
    bm = np.zeros((6, 5, 1), np.int8)
    bm[0, 0, 0] = 1
    ib = ImmersedBoundary(torch.zeros((6, 5, 1, 1)), 0.1, 1)
    ib.bitmask[...] = bm
    with pytest.raises(ValueError):
        ib.finalize()
    with pytest.raises(ValueError):
        O.immersed_boundary(u, bm, np.ones((1, 1), np.int8), np.zeros((1, 1)), 0.1)
    bm2 = np.zeros((6, 5, 1), np.int8)
    bm2[-1, 4, 0] = 1                 # +X neighbour after the array's end
    ib3 = ImmersedBoundary(torch.zeros((6, 5, 1, 1)), 0.1, 1)
    ib3.bitmask[...] = bm2
    with pytest.raises(ValueError):
        ib3.finalize()
    with pytest.raises(ValueError):
        O.immersed_boundary(u, bm2, np.ones((1, 1), np.int8), np.zeros((1, 1)), 0.1)
    bm4 = np.zeros((6, 5, 1), np.int8)
    bm4[2, 0, 0] = 1
    ib4 = ImmersedBoundary(torch.zeros((6, 5, 1, 1)), 0.1, 1)
    ib4.bitmask[...] = bm4
    ib4.finalize()
    out = O.immersed_boundary(u + 1.0, bm4, np.ones((1, 1), np.int8), np.full((1, 1), 2.0), 0.1)
    assert out[2, 0, 0, 0] == 2.0 and len(ib4.solid_boundary) == 1
