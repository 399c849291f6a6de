"""GPU parity of the Runge-Kutta integrators (reference
pm/time_step/rungeKatta.py) against the oracle's steps composed from the
NumPy FP32 restatement, bit for bit on the interior: a heat loop (RK1 and RK2,
eager and graph-replayed) and one step of the weakly compressible cavity of
examples/LDC/LDC_RK2.py (two fields, outer product, matrix divergence,
biharmonic damping, maps)."""
import numpy as np
import pytest
import torch

from conftest import same_bits
from helpers import ghost_free_mask
from oracle import np_oracle as O
from pde_module_amd import ForwardEuler, Laplacian
from pde_module_amd.deprecated import GridBoundary as LegacyGridBoundary
from pde_module_amd.FDM import Divergence, Grad, OuterProduct, functional, scalarVectorMult
from pde_module_amd.stencil import MapWise, element_ops
from pde_module_amd.time_step import RungeKatta1, RungeKatta2

pytestmark = pytest.mark.gpu


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _heat(dev, n=24, g=1):
    rng = np.random.default_rng(3)
    shape = (n + 2, n + 2, 1)
    u0 = rng.uniform(-1, 1, shape + (1,)).astype(np.float32)
    dx, alpha = 1.0 / (n - 1), 0.02
    dt = 0.4 * dx * dx / alpha
    ud = T(u0, dev)
    bcm = LegacyGridBoundary(ud, dx, g)
    bcm.dirichlet_BC("ALL", 0.0)
    bcm.dirichlet_BC("+X", 0.5)
    lap = Laplacian(1, dx, g)

    def f(t, u):
        return (lap(u, alpha),)

    def bc(t, u):
        return (bcm(u),)

    def of(t, u):
        return (O.laplacian(u, lap.stencil, np.float32(alpha), g),)

    def obc(t, u):
        return (O.bc_legacy(u, bcm.boundary_ijk_indices, bcm.boundary_interior,
                            bcm.boundary_type, bcm.boundary_value, dx),)

    return shape, g, u0, ud, dt, f, bc, of, obc


@pytest.mark.parametrize("order", [1, 2])
def test_heat_runge_kutta_matches_oracle(dev, order):
    shape, g, u0, ud, dt, f, bc, of, obc = _heat(dev)
    rk = (RungeKatta1 if order == 1 else RungeKatta2)([(1,)], f=f, bc=bc)
    ostep = O.rk1_step if order == 1 else O.rk2_step
    y, oy, t = ud, u0, 0.0
    for _ in range(5):
        (y,) = rk(t, dt, y)
        (oy,) = ostep(of, obc, t, dt, oy)
        t += dt
    assert same_bits(y.cpu().numpy(), oy, ghost_free_mask(shape, g, (1,)))


def test_rk2_graph_replay_matches_eager(dev):
    """One RK2 step captured once and replayed, with eager work and
    synchronisations between replays, equals the eager loop bit for bit."""
    shape, g, u0, ud, dt, f, bc, of, obc = _heat(dev)
    eager = RungeKatta2([(1,)], f=f, bc=bc)
    y = ud.clone()
    (y,) = eager(0.0, dt, y)
    ref = [y.clone()]
    for _ in range(3):
        (y,) = eager(0.0, dt, y)
        ref.append(y.clone())
    # captured: the first call (setup) is eager, like the reference example
    shape2, g2, u02, ud2, dt2, f2, bc2, of2, obc2 = _heat(dev)
    cap = RungeKatta2([(1,)], f=f2, bc=bc2)
    state = ud2.clone()
    (out,) = cap(0.0, dt2, state)
    state.copy_(out)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        (out,) = cap(0.0, dt2, state)
    m = torch.from_numpy(ghost_free_mask(shape, g, (1,))).to(dev)
    assert torch.equal(state[m].view(torch.int32), ref[0][m].view(torch.int32))
    for i in range(1, 4):
        graph.replay()
        torch.cuda.synchronize()
        state.copy_(out)
        _ = (state * 2).sum()          # eager work between replays
        torch.cuda.synchronize()
        assert torch.equal(state[m].view(torch.int32), ref[i][m].view(torch.int32)), i


def test_weakly_compressible_cavity_rk2_step_matches_oracle(dev):
    """One RK2 step of examples/LDC/LDC_RK2.py:101-144 (momentum m, density
    rho; u = m / rho; outer-product convection, matrix divergence, biharmonic
    damping, pressure c^2 (rho - rho0)).  The damping Laplacian reads the
    ghosts of the viscous term's output, which no kernel writes; they are
    zeroed once on both sides."""
    n, g = 17, 1
    dx = 1.0 / (n - 1)
    nu, rho0, U, M = 0.01, 1.0, 1.0, 0.1
    cs = U / M
    c2 = cs * cs
    dt = min(0.8 * dx * dx / nu, 0.8 * dx / cs, 0.8 * dx / (cs + U))
    eps = -0.01 * (np.exp(4 * np.log(dx) - np.log(dt)) / nu)
    shape = (n + 2, n + 2, 1)
    rng = np.random.default_rng(9)
    u0 = rng.uniform(-0.2, 0.2, shape + (2,)).astype(np.float32)
    r0 = rng.uniform(0.9, 1.1, shape + (1,)).astype(np.float32)
    ud, rd = T(u0, dev), T(r0, dev)
    u_BC = LegacyGridBoundary(ud, dx, g)
    u_BC.dirichlet_BC("ALL", 0.0)
    u_BC.dirichlet_BC("+Y", 1.0, 0)
    rho_BC = LegacyGridBoundary(rd, dx, g)
    rho_BC.vonNeumann_BC("ALL", 0.0)
    u_lapl = Laplacian(2, dx, g)
    u_outer = OuterProduct(2, 2)
    u_conv_div = Divergence((2, 2), ud.shape, dx, g)
    u_damping = Laplacian(2, dx, g)
    p_grad = Grad(1, rd.shape, dx, g)
    momentum_div = Divergence(2, ud.shape, dx, g)
    momentum = scalarVectorMult(2)
    get_p = MapWise(element_ops.shifted_scale)
    m_div_rho = MapWise("divide_by_scalar")
    uF = torch.empty_like(ud)

    def bc(t, m, rho):
        u = m_div_rho(m, rho)
        u_fix = u_BC(u)
        rho_fix = rho_BC(rho)
        return momentum(rho_fix, u_fix), rho_fix

    def f(t, m, rho):
        u = m_div_rho(m, rho)
        m_div = momentum_div(m, -1.0)
        u_diff = u_lapl(u, nu)
        if not hasattr(f, "zeroed"):
            gm = torch.from_numpy(~ghost_free_mask(shape, g, (2,))).to(dev)
            u_diff[gm] = 0.0
            f.zeroed = True
        u_conv = u_conv_div(u_outer(m, u))
        u_damp = u_damping(u_diff, eps)
        dp = p_grad(get_p(rho, c2, rho0), -1.0)
        functional.linear_combination(uF, (1, dp), (-1, u_conv), (1, u_diff), (1, u_damp))
        return uF, m_div

    lw, cw = O.laplacian_weights(dx), O.central_weights(dx)
    lb = lambda bcm, a: O.bc_legacy(a, bcm.boundary_ijk_indices, bcm.boundary_interior,
                                    bcm.boundary_type, bcm.boundary_value, dx)

    def obc(t, m, rho):
        u = O.divide_by_scalar(m, rho)
        uf, rf = lb(u_BC, u), lb(rho_BC, rho)
        return O.scalar_vector_multiply(rf, uf), rf

    def of(t, m, rho):
        u = O.divide_by_scalar(m, rho)
        m_div = O.divergence(m, cw, np.float32(-1.0), g)
        u_diff = O.laplacian(u, lw, np.float32(nu), g)
        u_diff = np.where(ghost_free_mask(shape, g, (2,)), u_diff, np.float32(0.0))
        u_conv = O.divergence(O.outer(m, u), cw, np.float32(1.0), g, matrix_in=True)
        u_damp = O.laplacian(u_diff, lw, np.float32(eps), g)
        dp = O.grad(O.shifted_scale(rho, c2, rho0), cw, np.float32(-1.0), g)
        return O.linear_combination((1, dp), (-1, u_conv), (1, u_diff), (1, u_damp)), m_div

    rk = RungeKatta2([(2,), (1,)], f=f, bc=bc)
    m = momentum(rd, ud)
    gm, gr = rk(0.0, dt, m, rd)
    om, orr = O.rk2_step(of, obc, 0.0, dt, O.scalar_vector_multiply(r0, u0), r0)
    assert same_bits(gm.cpu().numpy(), om, ghost_free_mask(shape, g, (2,)))
    assert same_bits(gr.cpu().numpy(), orr, ghost_free_mask(shape, g, (1,)))
    assert np.isfinite(om[ghost_free_mask(shape, g, (2,))]).all()
