"""GPU parity of the staircase immersed boundary against the oracle, bit for
bit over the whole field: 2D slot geometry of examples/wave2d/Interference.py
and a 3D sphere, Dirichlet / Neumann per component, out of place and in
place."""
import numpy as np
import pytest
import torch

from conftest import same_bits
from oracle import np_oracle as O
from pde_module_amd.FDM import ImmersedBoundary
from pde_module_amd.mesh import UniformGridMesh

pytestmark = pytest.mark.gpu


def _slot(W, L):
    def slot(x, y, z):
        xb = (x >= 0.3 * W) & (x <= 0.35 * W)
        yb = (y >= 0.58 * L) | ((y <= 0.52 * L) & (y >= 0.48 * L)) | (y <= 0.42 * L)
        return xb & yb
    return slot


def _sphere(c, r):
    return lambda x, y, z: (x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2 < r * r


CASES = [("slot", (61, 31, 1), 1, 2.0 / 30), ("slot", (41, 21, 1), 2, 0.05),
         ("sphere", (14, 14, 14), 3, 1.0 / 11)]


@pytest.mark.parametrize("geom,nodes,C,dx", CASES)
def test_immersed_boundary_matches_oracle(dev, geom, nodes, C, dx):
    print(f"immersed {geom} {nodes} C={C}", flush=True)
    grid = UniformGridMesh(dx=dx, nodes_per_axis=nodes, origin=(0.0, 0.0, 0.0), ghost_cells=1)
    shape = tuple(grid.nodal_grid.shape[:3])
    rng = np.random.default_rng(5)
    u = rng.uniform(-1, 1, shape + (C,)).astype(np.float32)
    ud = torch.from_numpy(u).to(dev)
    ib = ImmersedBoundary(ud, dx, 1)
    if geom == "slot":
        W, L = (nodes[0] - 1) * dx, (nodes[1] - 1) * dx
        ib.from_bool_func(_slot(W, L), grid.meshgrid)
    else:
        ib.from_bool_func(_sphere(0.5, 0.3), grid.meshgrid)
    ib.finalize()
    assert len(ib.solid_boundary) > 0
    ib.dirichlet_BC("ALL", 0.0)
    if C > 1:
        ib.vonNeumann_BC("ALL", 0.75, 1)
    if C > 2:
        ib.dirichlet_BC("ALL", -2.0, 2)
    got = ib(ud, 4.5).cpu().numpy()
    want = O.immersed_boundary(u, ib.bitmask, ib.boundary_type, ib.boundary_value, dx, 4.5)
    assert same_bits(got, want)
    assert same_bits(ud.cpu().numpy(), u)          # input untouched
    w = ud.clone()
    ib.apply_inplace(w, 4.5)
    assert same_bits(w.cpu().numpy(), want)
