"""GPU parity of the node-wise glue (element ops, maps, fused array-operator
chains) against the NumPy FP32 oracle, bit for bit, through the C ABI and the
module API; and one step of the reference's artificial-compressibility
lid-driven cavity loop (examples/LDC/LDC_No_Graph.py:83-92) built from the
modules, against the oracle composed the same way."""
import numpy as np
import pytest
import torch

from conftest import same_bits
from helpers import ghost_free_mask
from oracle import np_oracle as O
import pde_module_amd._native as N
from pde_module_amd import ForwardEuler
from pde_module_amd.deprecated import GridBoundary as LegacyGridBoundary
from pde_module_amd.FDM import (Divergence, Grad, Laplacian, OuterProduct, functional,
                                scalarVectorMult)
from pde_module_amd.stencil import ElementWise, MapWise, element_ops

pytestmark = pytest.mark.gpu


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _r(shape, seed, lo=-1.0, hi=1.0):
    return np.random.default_rng(seed).uniform(lo, hi, shape).astype(np.float32)


GRIDS = [(13, 11, 1), (5, 6, 7), (1, 1, 3)]


@pytest.mark.parametrize("grid", GRIDS)
def test_element_ops_match_oracle(dev, grid):
    print(f"element ops {grid}", flush=True)
    a, b = _r(grid + (2,), 1), _r(grid + (3,), 2)
    s = _r(grid + (1,), 3, 0.5, 2.0)
    m = _r(grid + (3, 2), 4)
    v2 = _r(grid + (2,), 5)
    outer = ElementWise(element_ops.outer)
    got = outer(T(a, dev), T(b, dev))
    assert got.shape == grid + (2, 3)
    assert same_bits(got.cpu().numpy(), O.outer(a, b))
    assert outer(T(a, dev), T(b, dev)) is got
    mv = ElementWise("matvec")(T(m, dev), T(v2, dev))
    assert same_bits(mv.cpu().numpy(), O.matvec(m, v2))
    sm = ElementWise("scalar_vector_multiply")(T(s, dev), T(m, dev))
    assert same_bits(sm.cpu().numpy(), O.scalar_vector_multiply(s, m))
    d = MapWise("divide_by_scalar")(T(b, dev), T(s, dev))
    assert same_bits(d.cpu().numpy(), O.divide_by_scalar(b, s))
    p = MapWise(element_ops.shifted_scale)(T(s, dev), 100.0, 1.0)
    assert same_bits(p.cpu().numpy(), O.shifted_scale(s, 100.0, 1.0))
    p1 = MapWise(element_ops.component_shifted_scale(1))(T(b, dev), -2.5, 0.25)
    assert same_bits(p1.cpu().numpy(), O.shifted_scale(b, -2.5, 0.25, comp=1))
    ad = ElementWise("add")(T(a, dev), T(v2, dev))
    sb = ElementWise("sub")(T(a, dev), T(v2, dev))
    assert same_bits(ad.cpu().numpy(), a + v2) and same_bits(sb.cpu().numpy(), a - v2)


@pytest.mark.parametrize("n", [1, 3, 4, 5, 1027])
def test_linear_combination_and_aliasing(dev, n):
    xs = [_r((n, 1, 1, 1), 10 + i) for i in range(4)]
    xd = [T(x, dev) for x in xs]
    out = torch.empty_like(xd[0])
    functional.linear_combination(out, (1, xd[0]), (-1, xd[1]), (1, xd[2]), (1, xd[3]))
    assert same_bits(out.cpu().numpy(), ((xs[0] - xs[1]) + xs[2]) + xs[3])
    functional.linear_combination(out, (0.5, xd[0]), (-3.0, xd[1]))
    assert same_bits(out.cpu().numpy(), O.linear_combination((0.5, xs[0]), (-3.0, xs[1])))
    # in place: out is the first term
    functional.linear_combination(xd[0], (1, xd[0]), (2.0, xd[2]))
    assert same_bits(xd[0].cpu().numpy(), O.linear_combination((1, xs[0]), (2.0, xs[2])))


def test_fdm_element_modules(dev):
    grid = (9, 8, 1)
    m, u = _r(grid + (2,), 20), _r(grid + (2,), 21)
    rho = _r(grid + (1,), 22, 0.5, 1.5)
    op = OuterProduct(2, 2)
    mu = op(T(m, dev), T(u, dev))
    assert same_bits(mu.cpu().numpy(), O.outer(m, u))
    mom = scalarVectorMult(2)
    r = mom(T(rho, dev), T(u, dev))
    assert same_bits(r.cpu().numpy(), O.scalar_vector_multiply(rho, u))
    out = torch.empty(grid + (2, 2), device=dev)
    functional.elementWiseOp("outer", T(m, dev), T(u, dev), out)
    assert same_bits(out.cpu().numpy(), O.outer(m, u))
    with pytest.raises(AssertionError):
        OuterProduct(2, 3)(T(m, dev), T(u, dev))
    with pytest.raises(TypeError):
        ElementWise("no_such_op")
    with pytest.raises(TypeError):
        MapWise(None)


def test_acm_cavity_step_matches_oracle(dev, corc):
    """One step of the ACM lid-driven cavity loop: BCs, Laplacian, Grad
    (vec2 -> mat22), Divergence, du*u, pressure gradient, the chain
    dp - u_conv + u_diff fused in one pass, and two ForwardEuler updates."""
    n, g = 21, 1
    dx = 1.0 / (n - 1)
    nu, beta = 0.01, 0.3
    dt = min(0.5 * dx * dx / nu, 0.5 * dx / (1 + 1 / beta))
    shape = (n + 2, n + 2, 1)
    u0, p0 = _r(shape + (2,), 30, -0.1, 0.1), _r(shape + (1,), 31, -0.1, 0.1)
    ud, pd = T(u0, dev), T(p0, dev)
    u_bc = LegacyGridBoundary(ud, dx, g)
    u_bc.dirichlet_BC("ALL", 0.0)
    u_bc.dirichlet_BC("+Y", 1.0, 0)
    p_bc = LegacyGridBoundary(pd, dx, g)
    p_bc.vonNeumann_BC("ALL", 0.0)
    u_lapl, u_grad = Laplacian(2, dx, g), Grad(2, ud.shape, dx, ghost_cells=g)
    u_div, p_grad = Divergence(2, ud.shape, dx, g), Grad(1, pd.shape, dx, g)
    u_step, p_step = ForwardEuler(), ForwardEuler()
    conv = ElementWise("matvec")

    u_fix = u_bc(ud)
    u_diff = u_lapl(u_fix, nu)
    du = u_grad(u_fix)
    u_incomp = u_div(u_fix, -beta)
    u_conv = conv(du, ud)
    p_fix = p_bc(pd)
    dp = p_grad(p_fix, -1.0)
    u_F = torch.empty_like(ud)
    functional.linear_combination(u_F, (1, dp), (-1, u_conv), (1, u_diff))
    u_next = u_step(u_fix, u_F, dt)
    p_next = p_step(p_fix, u_incomp, dt)

    ou = corc.bc_legacy(u0, u_bc.boundary_ijk_indices, u_bc.boundary_interior,
                        u_bc.boundary_type, u_bc.boundary_value, dx)
    op_ = corc.bc_legacy(p0, p_bc.boundary_ijk_indices, p_bc.boundary_interior,
                         p_bc.boundary_type, p_bc.boundary_value, dx)
    lw, cw = O.laplacian_weights(dx), O.central_weights(dx)
    o_diff = O.laplacian(ou, lw, np.float32(nu), g)
    o_du = O.grad(ou, cw, np.float32(1.0), g)
    o_inc = O.divergence(ou, cw, np.float32(-beta), g)
    o_conv = O.matvec(o_du, u0)
    o_dp = O.grad(op_, cw, np.float32(-1.0), g)
    o_F = O.linear_combination((1, o_dp), (-1, o_conv), (1, o_diff))
    o_un = O.forward_euler(ou, o_F, dt)
    o_pn = O.forward_euler(op_, o_inc, dt)
    mu = ghost_free_mask(shape, g, (2,))
    mp = ghost_free_mask(shape, g, (1,))
    assert same_bits(u_conv.cpu().numpy(), o_conv, mu)
    assert same_bits(u_next.cpu().numpy(), o_un, mu)
    assert same_bits(p_next.cpu().numpy(), o_pn, mp)


def test_function_maps_and_element_ops(dev):
    """User functions in MapWise / ElementWise (the cylinder example's
    ``get_p_op`` and ``m / rho[0]`` maps, examples/Cylinder/Cylinder_Transient.py:170-176,
    written over whole fields) run on the device into the modules' fixed
    outputs: equal to the oracle's restatements of the same maps (the
    subtract-multiply bit for bit; torch's device division within one ulp),
    the same output object on every call, and replayed from a captured graph
    with the same bits as eagerly."""
    print("function maps and element ops", flush=True)
    grid = (37, 29, 1)
    rho = _r(grid + (1,), 7, 0.5, 2.0)
    m = _r(grid + (2,), 8)
    get_p = MapWise(lambda r, c2, density: c2 * (r - density))
    p = get_p(T(rho, dev), 100.0, 1.0)
    assert same_bits(p.cpu().numpy(), O.shifted_scale(rho, 100.0, 1.0))
    assert get_p(T(rho, dev), 100.0, 1.0) is p
    get_u = MapWise(lambda mm, r: mm / r[..., 0:1])
    u = get_u(T(m, dev), T(rho, dev)).cpu().numpy()
    want = O.divide_by_scalar(m, rho)
    assert np.all(np.abs(u.view(np.int32).astype(np.int64) - want.view(np.int32)) <= 1)
    prod = ElementWise(lambda a, b: a * b[..., 0:1] + 1.0)
    got = prod(T(m, dev), T(rho, dev)).cpu().numpy()
    assert got.shape == grid + (2,)
    assert same_bits(got, m * rho + np.float32(1.0))
    # an output shape other than A's, given explicitly
    norm = ElementWise(lambda a, b: (a * a).sum(-1, keepdim=True) + b, output_dtype=(1,))
    got = norm(T(m, dev), T(rho, dev)).cpu().numpy()
    assert np.allclose(got, (m * m).sum(-1, keepdims=True) + rho, rtol=1e-6)
    # graph replay: the maps' torch operations are captured with their output
    a, b = T(m, dev), T(rho, dev)
    eager = prod(a, b).clone()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        out = prod(a, b)
    out.zero_()
    for _ in range(3):
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, eager)
