"""GPU parity of the absorbing layers (FarField sponge, ViscousDampingLayer)
against the oracle's node-by-node restatement, bit for bit over the whole
field (nodes outside the layer stay 0), and one damped wave step of the
reference's interference example (examples/wave2d/Interference.py:139-146)."""
import numpy as np
import pytest
import torch

from conftest import same_bits
from helpers import ghost_free_mask
from oracle import np_oracle as O
from pde_module_amd import ForwardEuler, Laplacian
from pde_module_amd.deprecated import GridBoundary as LegacyGridBoundary
from pde_module_amd.FDM import FarField, ViscousDampingLayer, functional

pytestmark = pytest.mark.gpu


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.mark.parametrize("shape,C,n,exclude", [((24, 17, 1), 1, 4, ()), ((19, 1, 1), 2, 3, ("+X",)),
                                               ((12, 10, 9), 3, 2, ("-Y", "+Z")),
                                               ((30, 22, 1), (2, 2), 6, ("-X",))])
def test_far_field_and_damping_match_oracle(dev, shape, C, n, exclude):
    print(f"layers shape={shape} C={C} n={n} exclude={exclude}", flush=True)
    elem = (C,) if isinstance(C, int) else C
    rng = np.random.default_rng(21)
    u = rng.uniform(-1, 1, shape + elem).astype(np.float32)
    ffv = rng.uniform(-0.5, 0.5, int(np.prod(elem))).astype(np.float32)
    ff = FarField(C, n, 2.5, shape, 0.05, 1)
    ff.exclude(*exclude)
    got = ff(T(u, dev), ffv, 7.5)
    want = O.damping_layer(u, shape, n, 2.5, ff.get_sponge_array(), 7.5, farfield=ffv)
    assert same_bits(got.cpu().numpy(), want)
    assert ff(T(u, dev), ffv, 7.5) is got
    vd = ViscousDampingLayer(C, n, shape, 0.05, 1)
    vd.exclude(*exclude)
    got = vd(T(u, dev), 3.25)
    want = O.damping_layer(u, shape, n, 2.0, vd.get_damping_array(), 3.25)
    assert same_bits(got.cpu().numpy(), want)


def test_damped_interference_step_matches_oracle(dev, corc):
    """u3 = BC(u); stencil = lapl(u3, c2); u_far = damp(v, C_max);
    v_next = v_step(v, stencil + u_far, dt); u_next = u_step(u3, v, dt),
    three steps with the reference's aliasing (outputs fed back, so from the
    second step on u_step reads the updated v)."""
    n, m, g = 31, 2, 1
    L = 2.0
    dx = L / (n - 1)
    c = 0.1
    c2 = c * c
    dt = 0.3 * float(1 / (c * np.sqrt(2.0 / dx ** 2)))
    shape = (n * m + 2 * g, n + 2 * g, 1)
    rng = np.random.default_rng(2)
    u0 = rng.uniform(-1, 1, shape + (1,)).astype(np.float32)
    v0 = np.zeros_like(u0)
    bc = LegacyGridBoundary(T(u0, dev), dx, g)
    bc.dirichlet_BC("ALL", 0.0)
    thick = 8
    damp = ViscousDampingLayer(1, thick, shape, dx, g)
    C_max = damp.calculate_C_max(c, dx * thick, damp.p, 1e-5)
    lapl = Laplacian(1, dx, g)
    u_step, v_step = ForwardEuler(), ForwardEuler()
    rhs = torch.empty_like(T(u0, dev))
    u, v = T(u0, dev), T(v0, dev)
    ou, ov = u0, v0
    pts = None
    for it in range(3):
        u3 = bc(u)
        stencil = lapl(u3, c2)
        u_far = damp(v, C_max)
        functional.linear_combination(rhs, (1, stencil), (1, u_far))
        v_next = v_step(v, rhs, dt)
        u_next = u_step(u3, v, dt)
        u, v = u_next, v_next
        pts = damp.get_damping_array() if pts is None else pts
        o3 = corc.bc_legacy(ou, bc.boundary_ijk_indices, bc.boundary_interior, bc.boundary_type,
                            bc.boundary_value, dx)
        ost = O.laplacian(o3, lapl.stencil, np.float32(c2), g)
        ofar = O.damping_layer(ov, shape, thick, 2.0, pts, C_max)
        ovn = O.forward_euler(ov, O.linear_combination((1, ost), (1, ofar)), dt)
        # from the second step on v is v_step's own output, updated in place
        # before u_step reads it (the reference loop's aliasing)
        oun = O.forward_euler(o3, ov if it == 0 else ovn, dt)
        ou, ov = oun, ovn
    mask = ghost_free_mask(shape, g, (1,))
    assert same_bits(u.cpu().numpy(), ou, mask)
    assert same_bits(v.cpu().numpy(), ov, mask)
