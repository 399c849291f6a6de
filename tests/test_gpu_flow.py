"""GPU checks of the projection cavity (pressure projection with Jacobi
sweeps): each step bit for bit against the oracle's composition, graph replay
equal to eager, the Jacobi / transposed mat-vec kernels against the oracle,
and dJ/du0 of the final kinetic energy (checkpointed rollout through the
transposed kernels) against central differences of the FP64 oracle loop."""
import numpy as np
import pytest
import torch

from conftest import same_bits
from helpers import ghost_free_mask
from oracle import np_oracle as O
import pde_module_amd._native as N
from pde_module_amd.flow import ProjectionCavity
from pde_module_amd.workloads import ns_cavity_u0

pytestmark = pytest.mark.gpu


def _oracle_loop(cav, u0, p0, steps, dtype):
    lw = O.laplacian_weights(cav.dx, dtype)
    cw = O.central_weights(cav.dx, dtype)
    lb = lambda bc, a: O.bc_legacy(a, bc.boundary_ijk_indices, bc.boundary_interior,
                                   bc.boundary_type, bc.boundary_value, cav.dx)
    u, p = u0.astype(dtype), p0.astype(dtype)
    h2 = dtype(cav.dx * cav.dx) if dtype == np.float64 else cav.h2
    for _ in range(steps):
        u, p = O.projection_step(u, p, lw, cw, cav.nu, cav.dt, h2, dtype(cav.inv),
                                 cav.sweeps, cav.g, lambda a: lb(cav.u_bc, a),
                                 lambda a: lb(cav.p_bc, a))
    return u, p


def test_projection_steps_match_oracle(dev):
    n = 34
    cav = ProjectionCavity(n, jacobi_sweeps=50, device=dev)
    u0 = ns_cavity_u0(0, n)
    p0 = np.zeros(cav.shape + (1,), np.float32)
    u, p = torch.from_numpy(u0).to(dev), torch.from_numpy(p0).to(dev)
    for _ in range(3):
        cav.step(u, p, u, p)
    ou, op = _oracle_loop(cav, u0, p0, 3, np.float32)
    assert same_bits(u.cpu().numpy(), ou, ghost_free_mask(cav.shape, 1, (2,)))
    assert same_bits(p.cpu().numpy(), op, ghost_free_mask(cav.shape, 1, (1,)))


def test_graph_replay_equals_eager(dev):
    n = 24
    cav = ProjectionCavity(n, jacobi_sweeps=10, device=dev)
    u0 = torch.from_numpy(ns_cavity_u0(1, n)).to(dev)
    p0 = torch.zeros(cav.shape + (1,), device=dev)
    ue, pe = cav.run(u0, p0, 12, use_graph=False)
    ug, pg = cav.run(u0, p0, 12, use_graph=True)
    torch.cuda.synchronize()
    assert torch.equal(ue.view(torch.int32), ug.view(torch.int32))
    assert torch.equal(pe.view(torch.int32), pg.view(torch.int32))


@pytest.mark.parametrize("shape", [(11, 9, 1), (7, 6, 8)])
def test_jacobi_and_transposes_match_oracle(dev, shape):
    rng = np.random.default_rng(4)
    p = rng.uniform(-1, 1, shape + (1,)).astype(np.float32)
    rhs = rng.uniform(-1, 1, shape + (1,)).astype(np.float32)
    nd = len(O.eligible_axes(shape))
    h2, inv = np.float32(0.013), np.float32(1.0 / (2 * nd))
    T = lambda a: torch.from_numpy(a).to(dev)
    pd, rd = T(p), T(rhs)              # kept alive across the launches
    out = torch.zeros_like(pd)
    N.call("pde_jacobi_f32", pd.data_ptr(), rd.data_ptr(), out.data_ptr(), *shape, 1, h2,
           inv, N.stream_of(out))
    m = ghost_free_mask(shape, 1, (1,))
    assert same_bits(out.cpu().numpy(), O.jacobi(p, rhs, 1, h2, inv), m)
    # transpose: dot-product test in FP64 of the FP32 results
    y = rng.uniform(-1, 1, shape + (1,)).astype(np.float32)
    gp, gr = torch.empty_like(out), torch.empty_like(out)
    yd = T(y)
    N.call("pde_jacobi_adjoint_f32", yd.data_ptr(), gp.data_ptr(), gr.data_ptr(), *shape, 1,
           h2, inv, N.stream_of(out))
    lhs = float(np.sum(np.where(m, out.cpu().numpy(), 0).astype(np.float64) * y))
    rhs_ = float(np.sum(gp.cpu().numpy().astype(np.float64) * p) +
                 np.sum(gr.cpu().numpy().astype(np.float64) * rhs))
    assert abs(lhs - rhs_) <= 1e-5 * (abs(lhs) + 1)
    # transposed mat-vec, 2 x 3 matrices
    M = rng.uniform(-1, 1, shape + (2, 3)).astype(np.float32)
    g = rng.uniform(-1, 1, shape + (2,)).astype(np.float32)
    gv = torch.empty(shape + (3,), device=dev)
    Md, gd = T(M), T(g)
    N.call("pde_matvec_t_f32", Md.data_ptr(), gd.data_ptr(), gv.data_ptr(),
           int(np.prod(shape)), 2, 3, N.stream_of(gv))
    want = M[..., 0, :] * g[..., 0:1] + M[..., 1, :] * g[..., 1:2]
    assert same_bits(gv.cpu().numpy(), want)


def test_kinetic_energy_gradient_matches_fp64_finite_differences(dev):
    n, steps = 18, 6
    cav = ProjectionCavity(n, jacobi_sweeps=20, dt=2e-3, device=dev)
    u0 = ns_cavity_u0(2, n) * 50.0                  # O(0.05) so convection matters
    p0 = np.zeros(cav.shape + (1,), np.float32)
    grads = []
    for K in (2, steps):
        x = torch.from_numpy(u0).to(dev).requires_grad_(True)
        J = cav.kinetic_energy_rollout(x, torch.from_numpy(p0).to(dev), steps, checkpoint_every=K)
        (gx,) = torch.autograd.grad(J, x)
        grads.append((float(J), gx.cpu().numpy().astype(np.float64)))
    assert np.array_equal(grads[0][1], grads[1][1])       # checkpoint-interval invariant
    J32, g32 = grads[0]
    mask = ghost_free_mask(cav.shape, 1, (2,))

    def J64(u):
        uN, _ = _oracle_loop(cav, u, p0, steps, np.float64)
        return 0.5 * cav.dx * cav.dx * float(np.sum(np.where(mask, uN, 0.0) ** 2))

    base = J64(u0.astype(np.float64))
    assert abs(J32 - base) <= 1e-5 * abs(base)
    rng = np.random.default_rng(7)
    for _ in range(3):
        d = np.where(mask, rng.standard_normal(u0.shape), 0.0)
        eps = 1e-3 * np.abs(u0).max()
        fd = (J64(u0 + eps * d) - J64(u0 - eps * d)) / (2 * eps)
        ad = float(np.sum(g32 * d))
        print(f"dJ/du0 . d: adjoint {ad:.8e}  fd {fd:.8e}", flush=True)
        assert abs(ad - fd) <= 1e-4 * abs(fd) + 1e-12


@pytest.mark.parametrize("shape", [(11, 9, 1), (7, 6, 8), (12, 1, 1)])
def test_folded_neumann_sweep_equals_boundary_then_sweep(dev, shape):
    """pde_jacobi_neumann0_f32 == legacy vonNeumann_BC('ALL', 0) + pde_jacobi_f32,
    bit for bit, including fields holding -0.0 next to the walls."""
    from pde_module_amd.deprecated import GridBoundary as LegacyGridBoundary
    rng = np.random.default_rng(8)
    p = rng.uniform(-1, 1, shape + (1,)).astype(np.float32)
    p[1, ...] = -0.0                      # signed zeros beside the -X wall
    p[-2, ...] = -0.0                     # and the +X wall
    rhs = rng.uniform(-1, 1, shape + (1,)).astype(np.float32)
    nd = len(O.eligible_axes(shape))
    h2, inv = np.float32(0.021), np.float32(1.0 / (2 * nd))
    pd, rd = torch.from_numpy(p).to(dev), torch.from_numpy(rhs).to(dev)
    bc = LegacyGridBoundary(pd, 0.1, 1)
    bc.vonNeumann_BC("ALL", 0.0)
    pf = bc(pd)
    a = torch.zeros_like(pd)
    b = torch.zeros_like(pd)
    N.call("pde_jacobi_f32", pf.data_ptr(), rd.data_ptr(), a.data_ptr(), *shape, 1, h2, inv,
           N.stream_of(a))
    N.call("pde_jacobi_neumann0_f32", pd.data_ptr(), rd.data_ptr(), b.data_ptr(), *shape, 1, h2,
           inv, N.stream_of(b))
    assert torch.equal(a.view(torch.int32), b.view(torch.int32))


@pytest.mark.parametrize("shape", [(11, 9, 1), (7, 6, 8), (12, 1, 1), (5, 5, 1)])
def test_folded_neumann_sweep_transpose(dev, shape):
    """pde_jacobi_neumann0_adjoint_f32 against the unfolded transpose
    (pde_jacobi_adjoint_f32, then the Neumann-0 boundary's transpose, grhs
    summed), and the dot-product identity <J(p, rhs), y> = <p, J^T_p y> +
    <rhs, J^T_rhs y> over interior p and rhs."""
    print(f"folded Jacobi transpose {shape}", flush=True)
    from pde_module_amd.deprecated import GridBoundary as LegacyGridBoundary
    rng = np.random.default_rng(21)
    nd = len(O.eligible_axes(shape))
    h2, inv = np.float32(0.021), np.float32(1.0 / (2 * nd))
    mask = ghost_free_mask(shape, 1, (1,))
    y = np.where(mask, rng.uniform(-1, 1, shape + (1,)), 0.0).astype(np.float32)
    g0 = rng.uniform(-1, 1, shape + (1,)).astype(np.float32)     # grhs accumulates onto g0
    yd = torch.from_numpy(y).to(dev)
    gp = torch.zeros_like(yd)
    grhs = torch.from_numpy(g0.copy()).to(dev)
    N.call("pde_jacobi_neumann0_adjoint_f32", yd.data_ptr(), gp.data_ptr(), grhs.data_ptr(),
           *shape, 1, h2, inv, N.stream_of(yd))
    # unfolded
    bc = LegacyGridBoundary(yd.clone(), 0.1, 1)
    bc.vonNeumann_BC("ALL", 0.0)
    bc._ensure_table(yd)
    a, ra = torch.empty_like(yd), torch.empty_like(yd)
    N.call("pde_jacobi_adjoint_f32", yd.data_ptr(), a.data_ptr(), ra.data_ptr(), *shape, 1, h2,
           inv, N.stream_of(yd))
    b = torch.empty_like(yd)
    bc.table.adjoint(a, b)
    got, want = gp.cpu().numpy(), b.cpu().numpy()
    assert np.all(got[~mask] == 0)
    np.testing.assert_allclose(got[mask], want[mask], rtol=2e-7, atol=2e-7)
    want_rhs = np.where(mask, g0 + ra.cpu().numpy(), g0)
    assert np.array_equal(grhs.cpu().numpy(), want_rhs.astype(np.float32))
    # dot products (FP64 sums of the FP32 kernels' outputs)
    p = np.where(mask, rng.uniform(-1, 1, shape + (1,)), 0.0).astype(np.float32)
    r = np.where(mask, rng.uniform(-1, 1, shape + (1,)), 0.0).astype(np.float32)
    out = torch.zeros_like(yd)
    pd, rd = torch.from_numpy(p).to(dev), torch.from_numpy(r).to(dev)   # alive past the launch
    N.call("pde_jacobi_neumann0_f32", pd.data_ptr(), rd.data_ptr(), out.data_ptr(), *shape, 1,
           h2, inv, N.stream_of(out))
    torch.cuda.synchronize()
    lhs = float(np.sum(out.cpu().numpy().astype(np.float64) * y))
    gr = grhs.cpu().numpy().astype(np.float64) - g0
    rhs_ = float(np.sum(p.astype(np.float64) * got) + np.sum(np.where(mask, r, 0.0) * gr))
    assert abs(lhs - rhs_) <= 1e-5 * max(1.0, abs(lhs))


def test_cavity_step_transpose_matches_autograd(dev):
    """The hand-written transposed rollout (CavityRollout) against torch
    autograd over the differentiable operators, for dJ/du0 and dJ/dp0; every
    storage mode and graph replay give the same bits."""
    n, steps = 18, 5
    cav = ProjectionCavity(n, jacobi_sweeps=12, dt=2e-3, device=dev)
    u0 = torch.from_numpy(ns_cavity_u0(4, n) * 50.0).to(dev)
    rng = np.random.default_rng(5)
    p0 = torch.from_numpy(rng.uniform(-1e-3, 1e-3, cav.shape + (1,)).astype(np.float32)).to(dev)

    def grads(fn, **kw):
        x, q = u0.clone().requires_grad_(True), p0.clone().requires_grad_(True)
        J = fn(x, q, steps, **kw)
        gx, gq = torch.autograd.grad(J, (x, q))
        return float(J), gx.cpu().numpy(), gq.cpu().numpy()

    Ja, gua, gpa = grads(cav.kinetic_energy_rollout_autograd, checkpoint_every=2)
    runs = [grads(cav.kinetic_energy_rollout, checkpoint_every=None),
            grads(cav.kinetic_energy_rollout, checkpoint_every=2),
            grads(cav.kinetic_energy_rollout, checkpoint_every=None, use_graph=True),
            grads(cav.kinetic_energy_rollout, checkpoint_every=3, use_graph=True)]
    J, gu, gp = runs[0]
    for Jr, gur, gpr in runs[1:]:
        assert Jr == J
        assert same_bits(gur, gu) and same_bits(gpr, gp)
    assert J == Ja
    # one rollout object run twice, graphs replayed on the buffers it kept
    from pde_module_amd.flow import CavityRollout
    ro = CavityRollout(cav, steps, checkpoint_every=2, use_graph=True)
    for _ in range(2):
        uN, _ = ro.forward_states(u0, p0)
        g_u, g_p = ro.backward_states(float(cav.dx * cav.dx) * uN * cav.mask, torch.zeros_like(p0))
        assert same_bits(g_u.cpu().numpy(), gu) and same_bits(g_p.cpu().numpy(), gp)
    mu = ghost_free_mask(cav.shape, 1, (2,))
    mp = ghost_free_mask(cav.shape, 1, (1,))
    scale_u, scale_p = np.abs(gua[mu]).max(), np.abs(gpa[mp]).max()
    print(f"|dJ/du0| {scale_u:.3e}  |dJ/dp0| {scale_p:.3e}", flush=True)
    np.testing.assert_allclose(gu, gua, rtol=0, atol=1e-5 * scale_u)
    np.testing.assert_allclose(gp, gpa, rtol=0, atol=1e-5 * max(scale_p, 1e-30))


@pytest.mark.parametrize("sweeps", [12, 13])
def test_cavity_batch_matches_single_instances(dev, sweeps):
    """ProjectionCavityBatch (sweeps of every instance in one launch) against
    each instance stepped alone: the same bits, forward (eager and graph) and
    dJ/du0, dJ/dp0 of the summed kinetic energy."""
    from pde_module_amd.flow import ProjectionCavityBatch
    print(f"cavity batch sweeps={sweeps}", flush=True)
    n, B, steps = 18, 3, 4
    kw = dict(jacobi_sweeps=sweeps, dt=2e-3, device=dev)
    bat = ProjectionCavityBatch(n, B, **kw)
    one = ProjectionCavity(n, **kw)
    rng = np.random.default_rng(17)
    u0 = torch.from_numpy(np.stack([ns_cavity_u0(s, n) * 50.0 for s in range(B)])).to(dev)
    p0 = torch.from_numpy(rng.uniform(-1e-3, 1e-3, (B,) + one.shape + (1,)).astype(np.float32)).to(dev)
    ue, pe = bat.run(u0, p0, steps, use_graph=False)
    ug, pg = bat.run(u0, p0, steps, use_graph=True)
    for i in range(B):
        ui, pi = one.run(u0[i], p0[i], steps, use_graph=False)
        for got in ((ue, pe), (ug, pg)):
            assert same_bits(got[0][i].cpu().numpy(), ui.cpu().numpy())
            assert same_bits(got[1][i].cpu().numpy(), pi.cpu().numpy())
    x, q = u0.clone().requires_grad_(True), p0.clone().requires_grad_(True)
    J = bat.kinetic_energy_rollout(x, q, steps, checkpoint_every=3, use_graph=True)
    gx, gq = torch.autograd.grad(J, (x, q))
    for i in range(B):
        xi, qi = u0[i].clone().requires_grad_(True), p0[i].clone().requires_grad_(True)
        Ji = one.kinetic_energy_rollout(xi, qi, steps, checkpoint_every=None)
        gxi, gqi = torch.autograd.grad(Ji, (xi, qi))
        assert same_bits(gx[i].cpu().numpy(), gxi.cpu().numpy())
        assert same_bits(gq[i].cpu().numpy(), gqi.cpu().numpy())


def test_cavity_rejects_fields_of_another_shape(dev):
    """The kernels take their extents from the cavity's grid: a field of any
    other shape raises on the host before a launch."""
    from pde_module_amd.flow import CavityRollout, ProjectionCavityBatch
    cav = ProjectionCavity(10, jacobi_sweeps=4, dt=1e-3, device=dev)
    u = torch.zeros(cav.shape + (2,), device=dev)
    p = torch.zeros(cav.shape + (1,), device=dev)
    with pytest.raises(AssertionError):
        cav.step(u, torch.zeros(cav.shape + (2,), device=dev), u, p)
    with pytest.raises(AssertionError):
        cav.step(u[:-1], p[:-1], u[:-1], p[:-1])
    with pytest.raises(AssertionError):
        cav.step_adjoint(u, u, p, u, torch.zeros(8, 12, 1, 1, device=dev))
    bat = ProjectionCavityBatch(10, 2, jacobi_sweeps=4, dt=1e-3, device=dev)
    with pytest.raises(AssertionError):
        bat.step(u, p, u, p)                      # no batch axis
    ro = CavityRollout(cav, 2)
    with pytest.raises(RuntimeError):
        ro.backward_states(u, p)                  # no forward yet
    ro.forward_states(u, p)
    with pytest.raises(AssertionError):
        ro.backward_states(p, p)                  # would broadcast


def _folded_sweep_np(p, rhs, g, h2, inv):
    """FP32 restatement of the folded Neumann-0 sweep on a 2D (X, Y) grid of
    ghost width g, in the kernels' operation order (rows, then columns; low
    neighbour, then high)."""
    X, Y = p.shape
    out = np.zeros_like(p)
    f = np.float32
    for x in range(g, X - g):
        for k in range(g, Y - g):
            up, dn = p[x - 1, k], p[x + 1, k]
            zl, zr = p[x, k - 1], p[x, k + 1]
            a_lo = f(-0.0) + dn if x - 1 < g else up
            a_hi = f(0.0) + up if x + 1 >= X - g else dn
            b_lo = f(-0.0) + zr if k - 1 < g else zl
            b_hi = f(0.0) + zl if k + 1 >= Y - g else zr
            acc = f(0.0)
            for v in (a_lo, a_hi, b_lo, b_hi):
                acc = f(acc + v)
            out[x, k] = f(inv * f(acc - f(h2 * rhs[x, k])))
    return out


def _folded_sweep_adjoint_np(gout, grhs, g, h2, inv):
    X, Y = gout.shape
    gp = np.zeros_like(gout)
    grhs = grhs.copy()
    f = np.float32
    for x in range(g, X - g):
        for k in range(g, Y - g):
            acc = f(0.0)
            if x + 1 < X - g:
                acc = f(acc + gout[x + 1, k])
            if x - 1 >= g:
                acc = f(acc + gout[x - 1, k])
            if k + 1 < Y - g:
                acc = f(acc + gout[x, k + 1])
            if k - 1 >= g:
                acc = f(acc + gout[x, k - 1])
            v = f(inv * acc)
            if x - 1 >= g and x - 2 < g:
                v = f(v + f(inv * gout[x - 1, k]))
            if x + 1 < X - g and x + 2 >= X - g:
                v = f(v + f(inv * gout[x + 1, k]))
            if k - 1 >= g and k - 2 < g:
                v = f(v + f(inv * gout[x, k - 1]))
            if k + 1 < Y - g and k + 2 >= Y - g:
                v = f(v + f(inv * gout[x, k + 1]))
            gp[x, k] = v
            grhs[x, k] = f(grhs[x, k] + f(-h2 * f(inv * gout[x, k])))
    return gp, grhs


@pytest.mark.parametrize("shape,g,batch", [((70, 300, 1), 1, 1), ((1, 33, 261), 1, 1),
                                           ((23, 41, 1), 2, 1), ((19, 70, 1), 1, 3),
                                           ((4, 5, 1), 1, 2), ((9, 500, 1), 1, 2),
                                           ((250, 249, 1), 3, 1)])
def test_folded_sweeps_2d_tiles_match_restatement(dev, shape, g, batch):
    """The 2D row-tile sweep and transpose (single and batched entries) against
    the FP32 restatement above, bit for bit: multi-strip rows, a partial last
    chunk, ghost width 2, 2 interior nodes per axis; ghosts of the sweep's
    output untouched, the transpose's gp 0 on ghosts."""
    print(f"folded 2D tiles {shape} g={g} batch={batch}", flush=True)
    rng = np.random.default_rng(31)
    ext = [s for s in shape if s > 1]
    full = (batch,) + shape + (1,)
    h2, inv = np.float32(0.013), np.float32(0.25)
    p = rng.uniform(-1, 1, full).astype(np.float32)
    rhs = rng.uniform(-1, 1, full).astype(np.float32)
    sentinel = rng.uniform(5, 6, full).astype(np.float32)
    pd, rd = torch.from_numpy(p).to(dev), torch.from_numpy(rhs).to(dev)
    out = torch.from_numpy(sentinel.copy()).to(dev)
    gp = torch.empty_like(pd)
    grhs = torch.from_numpy(sentinel.copy()).to(dev)
    st = N.stream_of(pd)
    if batch == 1:
        N.call("pde_jacobi_neumann0_f32", pd.data_ptr(), rd.data_ptr(), out.data_ptr(), *shape, g,
               h2, inv, st)
        N.call("pde_jacobi_neumann0_adjoint_f32", pd.data_ptr(), gp.data_ptr(), grhs.data_ptr(),
               *shape, g, h2, inv, st)
    else:
        N.call("pde_jacobi_neumann0_batch_f32", pd.data_ptr(), rd.data_ptr(), out.data_ptr(),
               batch, *shape, g, h2, inv, st)
        N.call("pde_jacobi_neumann0_adjoint_batch_f32", pd.data_ptr(), gp.data_ptr(),
               grhs.data_ptr(), batch, *shape, g, h2, inv, st)
    # two sweeps in one launch == two single sweeps (ghosts of out untouched)
    two = torch.from_numpy(sentinel.copy()).to(dev)
    mid = torch.from_numpy(sentinel.copy()).to(dev)
    ref = torch.from_numpy(sentinel.copy()).to(dev)
    N.call("pde_jacobi_neumann0_x2_f32", pd.data_ptr(), rd.data_ptr(), two.data_ptr(), batch,
           *shape, g, h2, inv, st)
    N.call("pde_jacobi_neumann0_batch_f32", pd.data_ptr(), rd.data_ptr(), mid.data_ptr(), batch,
           *shape, g, h2, inv, st)
    N.call("pde_jacobi_neumann0_batch_f32", mid.data_ptr(), rd.data_ptr(), ref.data_ptr(), batch,
           *shape, g, h2, inv, st)
    assert same_bits(two.cpu().numpy(), ref.cpu().numpy())
    # two transposed sweeps in one launch == two single ones (gp and grhs)
    gp2, gr2 = torch.empty_like(pd), torch.from_numpy(sentinel.copy()).to(dev)
    gm, gp1, gr1 = torch.empty_like(pd), torch.empty_like(pd), torch.from_numpy(sentinel.copy()).to(dev)
    N.call("pde_jacobi_neumann0_adjoint_x2_f32", pd.data_ptr(), gp2.data_ptr(), gr2.data_ptr(),
           batch, *shape, g, h2, inv, st)
    N.call("pde_jacobi_neumann0_adjoint_batch_f32", pd.data_ptr(), gm.data_ptr(), gr1.data_ptr(),
           batch, *shape, g, h2, inv, st)
    N.call("pde_jacobi_neumann0_adjoint_batch_f32", gm.data_ptr(), gp1.data_ptr(), gr1.data_ptr(),
           batch, *shape, g, h2, inv, st)
    assert same_bits(gp2.cpu().numpy(), gp1.cpu().numpy())
    assert same_bits(gr2.cpu().numpy(), gr1.cpu().numpy())
    o, gpo, gro = (t.cpu().numpy().reshape((batch,) + tuple(ext)) for t in (out, gp, grhs))
    p2, r2, s2 = (a.reshape((batch,) + tuple(ext)) for a in (p, rhs, sentinel))
    for b in range(batch):
        want = _folded_sweep_np(p2[b], r2[b], g, h2, inv)
        inner = np.zeros(tuple(ext), bool)
        inner[g:-g, g:-g] = True
        want = np.where(inner, want, s2[b])
        assert same_bits(o[b], want)
        wgp, wgr = _folded_sweep_adjoint_np(p2[b], s2[b], g, h2, inv)
        assert same_bits(gpo[b], wgp)
        assert same_bits(gro[b], wgr)


@pytest.mark.parametrize("shape,g", [((20, 33, 1), 1), ((1, 70, 300), 1), ((15, 12, 1), 2)])
def test_cavity_predictor_fused_equals_separate_kernels(dev, shape, g):
    """pde_cavity_predict_f32 against grad + matvec + laplacian + lincomb +
    Euler on zero-ghost work buffers, every padded node bit for bit (signed
    zeros on the ghosts included)."""
    rng = np.random.default_rng(41)
    u = rng.uniform(-1, 1, shape + (2,)).astype(np.float32)
    u[0, ...] = -0.0                                   # signed zeros in a ghost layer
    ud = torch.from_numpy(u).to(dev)
    dx = 0.05
    from pde_module_amd.FDM.module.grad import central_weights
    wl, wc = O.laplacian_weights(dx), central_weights(dx)
    nu, dt = np.float32(0.01), np.float32(1e-3)
    z = lambda *e: torch.zeros(shape + e, device=dev)
    du, conv, diff, F, us = z(2, 2), z(2), z(2), z(2), z(2)
    st = N.stream_of(ud)
    N.call("pde_grad_f32", ud.data_ptr(), du.data_ptr(), *shape, 2, g, *wc, np.float32(1.0), st)
    N.call("pde_matvec_f32", du.data_ptr(), ud.data_ptr(), conv.data_ptr(),
           int(np.prod(shape)), 2, 2, st)
    N.call("pde_laplacian_f32", ud.data_ptr(), diff.data_ptr(), *shape, 2, g, *wl, nu, st)
    N.call("pde_lincomb_f32", conv.data_ptr(), diff.data_ptr(), None, None, np.float32(-1.0),
           np.float32(1.0), np.float32(0), np.float32(0), 2, F.data_ptr(), F.numel(), st)
    N.call("pde_forward_euler_f32", ud.data_ptr(), F.data_ptr(), us.data_ptr(), us.numel(), dt, st)
    fused = torch.full_like(us, 7.0)
    N.call("pde_cavity_predict_f32", ud.data_ptr(), fused.data_ptr(), *shape, g, *wl, *wc, nu, dt,
           st)
    assert same_bits(fused.cpu().numpy(), us.cpu().numpy())


@pytest.mark.parametrize("shape,g", [((20, 33, 1), 1), ((1, 70, 300), 1), ((15, 12, 1), 2)])
def test_cavity_correction_fused_equals_boundary_grad_euler(dev, shape, g):
    """pde_cavity_correct_f32 against legacy vonNeumann_BC('ALL', 0) on p, then
    pde_grad_f32 into a zero-ghost buffer, then Euler with -dt over every node:
    the same bits everywhere, signed zeros beside the walls included."""
    from pde_module_amd.deprecated import GridBoundary as LegacyGridBoundary
    from pde_module_amd.FDM.module.grad import central_weights
    rng = np.random.default_rng(43)
    ext = [s for s in shape if s > 1]
    p = rng.uniform(-1, 1, shape + (1,)).astype(np.float32)
    pv = p.reshape(ext)
    pv[g + 1, :] = -0.0                    # mirrors of the -X wall's ghosts
    pv[:, -g - 2] = -0.0                   # and of the +Y wall's
    us2 = rng.uniform(-1, 1, shape + (2,)).astype(np.float32)
    us2.reshape(ext + [2])[0, :, :] = -0.0
    pd, ud = torch.from_numpy(p).to(dev), torch.from_numpy(us2).to(dev)
    dx = 0.05
    wc = central_weights(dx)
    dt = np.float32(2e-3)
    bc = LegacyGridBoundary(pd.clone(), dx, g)
    bc.vonNeumann_BC("ALL", 0.0)
    pf = bc(pd)
    gp = torch.zeros(shape + (1, 2), device=dev)
    want = torch.empty_like(ud)
    st = N.stream_of(pd)
    N.call("pde_grad_f32", pf.data_ptr(), gp.data_ptr(), *shape, 1, g, *wc, np.float32(1.0), st)
    N.call("pde_forward_euler_f32", ud.data_ptr(), gp.data_ptr(), want.data_ptr(), want.numel(),
           np.float32(-dt), st)
    got = torch.full_like(ud, 9.0)
    N.call("pde_cavity_correct_f32", pd.data_ptr(), ud.data_ptr(), got.data_ptr(), *shape, g, *wc,
           dt, st)
    assert same_bits(got.cpu().numpy(), want.cpu().numpy())


@pytest.mark.parametrize("shape,g", [((20, 33, 1), 1), ((1, 70, 300), 1), ((15, 12, 1), 2)])
def test_cavity_adjoint_predictor_fused_equals_separate_kernels(dev, shape, g):
    """pde_cavity_adjoint_predict_f32 against the chain of separate transposed
    kernels (scale, scale, Laplacian^T, grad, outer, matvec^T, grad^T,
    lincomb) on every padded node, bit for bit."""
    from pde_module_amd.FDM.module.grad import central_weights
    rng = np.random.default_rng(47)
    gus = rng.uniform(-1, 1, shape + (2,)).astype(np.float32)
    u2 = rng.uniform(-1, 1, shape + (2,)).astype(np.float32)
    gd, ud = torch.from_numpy(gus).to(dev), torch.from_numpy(u2).to(dev)
    dx = 0.05
    wl, wc = O.laplacian_weights(dx), central_weights(dx)
    nu, dt = np.float32(0.01), np.float32(2e-3)
    nodes = int(np.prod(shape))
    z = lambda *e: torch.zeros(shape + e, device=dev)
    gF, gconv, glap, du, gdu, gmv, ggr, want = z(2), z(2), z(2), z(2, 2), z(2, 2), z(2), z(2), z(2)
    st = N.stream_of(gd)
    N.call("pde_scale_f32", gd.data_ptr(), gF.data_ptr(), gF.numel(), dt, st)
    N.call("pde_scale_f32", gF.data_ptr(), gconv.data_ptr(), gF.numel(), np.float32(-1.0), st)
    N.call("pde_laplacian_adjoint_f32", gF.data_ptr(), glap.data_ptr(), *shape, 2, g, *wl, nu, st)
    N.call("pde_grad_f32", ud.data_ptr(), du.data_ptr(), *shape, 2, g, *wc, np.float32(1.0), st)
    N.call("pde_outer_f32", gconv.data_ptr(), ud.data_ptr(), gdu.data_ptr(), nodes, 2, 2, st)
    N.call("pde_matvec_t_f32", du.data_ptr(), gconv.data_ptr(), gmv.data_ptr(), nodes, 2, 2, st)
    N.call("pde_grad_adjoint_f32", gdu.data_ptr(), ggr.data_ptr(), *shape, 2, g, *wc,
           np.float32(1.0), st)
    one = np.float32(1.0)
    N.call("pde_lincomb_f32", gd.data_ptr(), glap.data_ptr(), gmv.data_ptr(), ggr.data_ptr(), one,
           one, one, one, 4, want.data_ptr(), want.numel(), st)
    got = torch.full_like(want, 3.0)
    N.call("pde_cavity_adjoint_predict_f32", gd.data_ptr(), ud.data_ptr(), got.data_ptr(), *shape,
           g, *wl, *wc, nu, dt, st)
    assert same_bits(got.cpu().numpy(), want.cpu().numpy())


@pytest.mark.parametrize("shape,g", [((20, 33, 1), 1), ((1, 70, 300), 1), ((15, 12, 1), 2)])
def test_cavity_adjoint_correct_and_div_fused_equal_chains(dev, shape, g):
    """pde_cavity_adjoint_correct_f32 == scale(-dt) + grad transpose, and
    pde_cavity_adjoint_div_f32 == divergence transpose + lincomb, on every
    padded node, bit for bit."""
    from pde_module_amd.FDM.module.grad import central_weights
    rng = np.random.default_rng(53)
    wc = central_weights(0.05)
    dt = np.float32(2e-3)
    alpha = np.float32(1.0 / dt)
    gu = torch.from_numpy(rng.uniform(-1, 1, shape + (2,)).astype(np.float32)).to(dev)
    gr = torch.from_numpy(rng.uniform(-1, 1, shape + (1,)).astype(np.float32)).to(dev)
    st = N.stream_of(gu)
    ggp, want_p = torch.empty_like(gu), torch.empty_like(gr)
    N.call("pde_scale_f32", gu.data_ptr(), ggp.data_ptr(), gu.numel(), np.float32(-dt), st)
    N.call("pde_grad_adjoint_f32", ggp.data_ptr(), want_p.data_ptr(), *shape, 1, g, *wc,
           np.float32(1.0), st)
    got_p = torch.full_like(gr, 4.0)
    N.call("pde_cavity_adjoint_correct_f32", gu.data_ptr(), got_p.data_ptr(), *shape, g, *wc, dt,
           st)
    assert same_bits(got_p.cpu().numpy(), want_p.cpu().numpy())
    gdiv, want_u = torch.empty_like(gu), torch.empty_like(gu)
    N.call("pde_divergence_adjoint_f32", gr.data_ptr(), gdiv.data_ptr(), *shape, 1, g, *wc, alpha,
           st)
    N.call("pde_lincomb_f32", gu.data_ptr(), gdiv.data_ptr(), None, None, np.float32(1.0),
           np.float32(1.0), np.float32(0), np.float32(0), 2, want_u.data_ptr(), want_u.numel(), st)
    got_u = torch.full_like(gu, 4.0)
    N.call("pde_cavity_adjoint_div_f32", gr.data_ptr(), gu.data_ptr(), got_u.data_ptr(), *shape, g,
           *wc, alpha, st)
    assert same_bits(got_u.cpu().numpy(), want_u.cpu().numpy())


@pytest.mark.parametrize("n,sweeps,spl,batch", [
    (34, 23, 10, 1),     # one whole-grid tile, 3 sweeps left for the last launch
    (200, 21, 10, 1),    # 64-cell regions cut from the grid, 1 sweep left
    (130, 50, 7, 2),     # a batch of cut grids, 1 sweep left
    (18, 12, 5, 3),      # a batch of whole-grid tiles, a 2-sweep last launch
    (300, 40, 40, 1),    # 128-cell regions around 40-cell halos
    (40, 6, 10, 2),      # fewer sweeps than one launch takes
    (40, 1, 10, 1),      # a single sweep (one-sweep kernels, rhs gradient cleared first)
    (40, 0, 10, 1),      # no sweep at all
])
def test_multisweep_launches_equal_two_sweep_kernels(dev, n, sweeps, spl, batch):
    """Jacobi sweeps run ``spl`` per launch on LDS-resident tiles against the
    two-sweep row kernels: the same bits for three forward steps and for the
    transposed step (gradients of u, p and, through them, of rhs)."""
    from pde_module_amd.flow import ProjectionCavityBatch
    print(f"multisweep n={n} sweeps={sweeps} per_launch={spl} batch={batch}", flush=True)
    kw = dict(jacobi_sweeps=sweeps, dt=1e-3, device=dev)
    a = ProjectionCavityBatch(n, batch, sweeps_per_launch=spl, **kw)
    b = ProjectionCavityBatch(n, batch, sweeps_per_launch=2, **kw)
    rng = np.random.default_rng(n + sweeps)
    shp = (batch,) + a.cav.shape
    u0 = torch.from_numpy(np.stack([ns_cavity_u0(s, n) * 50.0 for s in range(batch)])).to(dev)
    p0 = torch.from_numpy(rng.uniform(-1e-3, 1e-3, shp + (1,)).astype(np.float32)).to(dev)
    ua, pa = a.run(u0, p0, 3, use_graph=False)
    ub, pb = b.run(u0, p0, 3, use_graph=False)
    assert same_bits(ua.cpu().numpy(), ub.cpu().numpy())
    assert same_bits(pa.cpu().numpy(), pb.cpu().numpy())
    gu_out = torch.from_numpy(rng.standard_normal(shp + (2,)).astype(np.float32)).to(dev)
    gp_out = torch.from_numpy(rng.standard_normal(shp + (1,)).astype(np.float32)).to(dev)
    got = []
    for c in (a, b):
        for rep in range(2):      # nothing of the first transposed step carries over
            gu, gp = torch.zeros_like(gu_out), torch.zeros_like(gp_out)
            c.step_adjoint(ua, gu_out, gp_out, gu, gp)
            got.append((gu.cpu().numpy(), gp.cpu().numpy()))
    for other in got[1:]:
        assert same_bits(got[0][0], other[0])
        assert same_bits(got[0][1], other[1])
    if batch == 1:
        # the single-instance class takes the same launches
        one = ProjectionCavity(n, sweeps_per_launch=spl, **kw)
        u1, p1 = one.run(u0[0], p0[0], 3, use_graph=True)
        assert same_bits(u1.cpu().numpy(), ub[0].cpu().numpy())
        assert same_bits(p1.cpu().numpy(), pb[0].cpu().numpy())
        gu, gp = torch.zeros_like(gu_out[0]), torch.zeros_like(gp_out[0])
        one.step_adjoint(ua[0], gu_out[0], gp_out[0], gu, gp)
        assert same_bits(gu.cpu().numpy(), got[0][0][0])
        assert same_bits(gp.cpu().numpy(), got[0][1][0])


def test_multisweep_at_the_cavity_workload_size(dev):
    """The 1024^2 cavity (50 sweeps, 10 per launch: 576 tiles of 64 x 64 nodes)
    against the two-sweep row kernels: two forward steps and a transposed step,
    bit for bit."""
    n = 1024
    print(f"multisweep at {n}^2", flush=True)
    kw = dict(jacobi_sweeps=50, device=dev)
    a = ProjectionCavity(n, sweeps_per_launch=10, **kw)
    b = ProjectionCavity(n, sweeps_per_launch=2, **kw)
    u0 = torch.from_numpy(ns_cavity_u0(0, n)).to(dev)
    rng = np.random.default_rng(5)
    p0 = torch.from_numpy(rng.uniform(-1e-3, 1e-3, a.shape + (1,)).astype(np.float32)).to(dev)
    ua, pa = a.run(u0, p0, 2, use_graph=False)
    ub, pb = b.run(u0, p0, 2, use_graph=False)
    assert torch.equal(ua.view(torch.int32), ub.view(torch.int32))
    assert torch.equal(pa.view(torch.int32), pb.view(torch.int32))
    gu_out = torch.from_numpy(rng.standard_normal(a.shape + (2,)).astype(np.float32)).to(dev)
    gp_out = torch.from_numpy(rng.standard_normal(a.shape + (1,)).astype(np.float32)).to(dev)
    got = []
    for c in (a, b):
        gu, gp = torch.zeros_like(gu_out), torch.zeros_like(gp_out)
        c.step_adjoint(ua, gu_out, gp_out, gu, gp)
        got.append((gu, gp))
    assert torch.equal(got[0][0].view(torch.int32), got[1][0].view(torch.int32))
    assert torch.equal(got[0][1].view(torch.int32), got[1][1].view(torch.int32))
