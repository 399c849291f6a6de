"""GPU parity of boundaries, ForwardEuler and whole time loops against the C
FP32 restatement (bit for bit on non-ghost nodes and on every written ghost),
plus the accumulated error of the FP32 heat loop against the FP64 oracle."""
import numpy as np
import pytest
import torch

from conftest import same_bits
from helpers import face_ghost_mask, ghost_free_mask
from oracle import np_oracle as O
from pde_module_amd import ForwardEuler, HeatStep, Laplacian, WaveStep
from pde_module_amd.deprecated import GridBoundary as LegacyGridBoundary
from pde_module_amd.FDM import GridBoundary
from pde_module_amd.FDM.boundary import get_ramp_func
from pde_module_amd.workloads import heat_2d_config

pytestmark = pytest.mark.gpu


def to_dev(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def oracle_groups(bc):
    out = []
    for fb in bc.func_groups.values():
        comps = fb.output_ids if isinstance(fb.output_ids, (list, tuple)) else [fb.output_ids]
        out.append(dict(ids=fb.face_ids, type=fb.boundary_type, comps=list(comps),
                        const=fb.func.constant, tmax=getattr(fb.func, "T_max", None)))
    return out


CASES = {
    "lid": ((34, 30, 1), 2, lambda b: (b.no_slip("ALL"), b.dirichlet_BC("+Y", 1.0, 0))),
    "neumann3d": ((12, 13, 14), 3, lambda b: (b.vonNeumann_BC("ALL", 0.5),
                                               b.dirichlet_BC("-Z", -1.0, [0, 2]))),
    "ramp": ((20, 21, 1), 1, lambda b: b.dirichlet_BC("ALL", get_ramp_func(3.0, 0.4))),
    "big": ((300, 260, 1), 2, lambda b: (b.impermeable("-X"), b.dirichlet_BC("+Y", 1.0, 0))),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_functional_boundary_matches_oracle(dev, corc, name):
    print(f"functional boundary {name}", flush=True)
    shape, C, cfg = CASES[name]
    rng = np.random.default_rng(1)
    u = rng.uniform(-1, 1, shape + (C,)).astype(np.float32)
    ud = to_dev(u, dev)
    bc = GridBoundary(ud, 0.1, 1)
    cfg(bc)
    for t in (0.0, 0.1, 1.0):
        got = bc(ud, t).cpu().numpy()
        want = corc.bc_functional(u, 1, oracle_groups(bc), 0.1, t=t)
        assert same_bits(got, want)
    # in place: same values on everything written, input elsewhere
    tt = torch.full((1,), 0.1, device=dev)
    uu = ud.clone()
    bc.apply_inplace(uu, tt)
    assert same_bits(uu.cpu().numpy(), corc.bc_functional(u, 1, oracle_groups(bc), 0.1, t=0.1))


@pytest.mark.parametrize("shape,C", [((40, 33, 1), 1), ((9, 10, 11), 3), ((600, 500, 1), 1),
                                     ((70, 70, 70), 1)])
def test_legacy_boundary_matches_oracle(dev, corc, shape, C):
    print(f"legacy boundary {shape} C={C}", flush=True)
    rng = np.random.default_rng(2)
    u = rng.uniform(-1, 1, shape + (C,)).astype(np.float32)
    ud = to_dev(u, dev)
    bc = LegacyGridBoundary(ud, 0.05, 1)
    bc.vonNeumann_BC("-X", 0.25)
    if C == 3:
        bc.impermeable("+Z")
    got = bc(ud).cpu().numpy()
    want = corc.bc_legacy(u, bc.boundary_ijk_indices, bc.boundary_interior, bc.boundary_type,
                          bc.boundary_value, 0.05)
    assert same_bits(got, want)
    uu = ud.clone()
    bc.apply_inplace(uu)   # one workgroup up to 16384 ops, gather/scatter beyond (70^3)
    assert same_bits(uu.cpu().numpy(), want)


def test_forward_euler_matches_oracle_and_aliases(dev, corc):
    rng = np.random.default_rng(3)
    for n in (1, 7, 1000, 1 << 20):
        a = rng.uniform(-1, 1, (n, 1, 1, 1)).astype(np.float32)
        k = rng.uniform(-1, 1, (n, 1, 1, 1)).astype(np.float32)
        e = ForwardEuler()
        got = e(to_dev(a, dev), to_dev(k, dev), 0.013).cpu().numpy()
        assert same_bits(got, corc.forward_euler(a, k, np.float32(0.013)))
    # in place through the module's own output (reference wave-loop aliasing)
    v = to_dev(a, dev)
    e = ForwardEuler()
    out = e(v, to_dev(k, dev), 0.5)
    out2 = e(out, to_dev(k, dev), 0.5)
    assert out2.data_ptr() == out.data_ptr()
    want = corc.forward_euler(corc.forward_euler(a, k, np.float32(0.5)), k, np.float32(0.5))
    assert same_bits(out2.cpu().numpy(), want)


def oracle_heat_loop(corc, u, lapw, alpha, dt, g, bc_fn, steps):
    """Reference heat loop on the C restatement; yields u after each step."""
    for _ in range(steps):
        u2 = bc_fn(u)
        L = corc.laplacian(u2, lapw, np.float32(alpha), g)
        u = corc.forward_euler(u2, L, np.float32(dt))
        yield u


def test_heat_loop_module_by_module_and_fused(dev, corc):
    """Heat workload (256^2, 8 seeded bumps, Dirichlet 0, dt = 0.2 dx^2/alpha),
    1000 steps: the reference-shaped module loop and the fused HeatStep both
    match the C restatement bit for bit at every step; the accumulated error
    against the FP64 oracle is reported."""
    cfg = heat_2d_config(seed=0)
    u0 = cfg.u0
    shape, g = u0.shape[:3], cfg.g
    ud = to_dev(u0, dev)
    bc = LegacyGridBoundary(ud, cfg.dx, g)
    bc.dirichlet_BC("ALL", 0.0)
    lap = Laplacian(1, cfg.dx, g)
    euler = ForwardEuler()
    bc2 = LegacyGridBoundary(ud, cfg.dx, g)
    bc2.dirichlet_BC("ALL", 0.0)
    fused = HeatStep(bc2, cfg.dx, g, cfg.alpha, cfg.dt)
    bc_fn = lambda x: corc.bc_legacy(x, bc.boundary_ijk_indices, bc.boundary_interior,
                                     bc.boundary_type, bc.boundary_value, cfg.dx)
    mask = ghost_free_mask(shape, g, (1,))
    u_mod, u_fus = ud, ud.clone()
    u64 = u0.astype(np.float64)
    ijk, _, inter = O.legacy_boundary_layout(shape, g)
    bt, bv = np.ones((len(ijk), 1), np.int8), np.zeros((len(ijk), 1))
    bc64 = lambda x: O.bc_legacy(x, ijk, inter, bt, bv, cfg.dx)
    w64 = O.laplacian_weights(cfg.dx, np.float64)
    errs = {}
    for step, want in enumerate(oracle_heat_loop(corc, u0, lap.stencil, cfg.alpha, cfg.dt, g,
                                                 bc_fn, 1000), start=1):
        u2 = bc(u_mod)
        u_mod = euler(u2, lap(u2, cfg.alpha), cfg.dt)
        u_fus = fused(u_fus)
        u64 = O.heat_step(u64, w64, cfg.alpha, g, cfg.dt, bc64)
        if step in (1, 2, 10, 100, 500, 1000):
            a = u_mod.cpu().numpy()
            b = u_fus.cpu().numpy()
            assert same_bits(a, want, mask), f"module loop differs at step {step}"
            assert same_bits(b, want, mask), f"fused step differs at step {step}"
            errs[step] = float(np.max(np.abs(a[mask] - u64[mask])) / np.max(np.abs(u64[mask])))
    print("heat 256^2 accumulated rel. error vs FP64:", errs, flush=True)
    assert errs[1000] < 1e-5


def test_wave_loop_module_by_module_and_fused(dev, corc):
    """Standing-wave loop (reference examples/wave2d/standingWave.py:70-76) with
    its in-place v aliasing, module by module and fused, vs the C restatement."""
    n, g = 65, 1
    dx = 1.0 / (n - 1)
    dt = 0.5 * dx
    xs = np.arange(-1, n + 1) * dx
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    u0 = (np.sin(2 * np.pi * X) * np.sin(2 * np.pi * Y)).astype(np.float32).reshape(n + 2, n + 2, 1, 1)
    v0 = np.zeros_like(u0)
    c2 = np.float32(1.0)
    ud, vd = to_dev(u0, dev), to_dev(v0, dev)
    bc = LegacyGridBoundary(ud, dx, g)
    bc.dirichlet_BC("ALL", 0.0)
    lap = Laplacian(1, dx, g)
    u_step, v_step = ForwardEuler(), ForwardEuler()
    bc2 = LegacyGridBoundary(ud, dx, g)
    bc2.dirichlet_BC("ALL", 0.0)
    fused = WaveStep(bc2, dx, g, float(c2), dt)
    bc_fn = lambda x: corc.bc_legacy(x, bc.boundary_ijk_indices, bc.boundary_interior,
                                     bc.boundary_type, bc.boundary_value, dx)
    mask = ghost_free_mask(u0.shape[:3], g, (1,))
    u, v = ud, vd
    fu, fv = ud.clone(), vd.clone()
    ou, ov = u0, v0
    for i in range(300):
        # reference loop
        u2 = bc(u)
        st = lap(u2, float(c2))
        v_next = v_step(v, st, dt)
        u_next = u_step(u2, v, dt)
        u, v = u_next, v_next
        fu, fv = fused(fu, fv)
        # oracle with the same aliasing
        ou2 = bc_fn(ou)
        L = corc.laplacian(ou2, lap.stencil, c2, g)
        nv = corc.forward_euler(ov, L, np.float32(dt))
        ou = corc.forward_euler(ou2, ov if i == 0 else nv, np.float32(dt))
        ov = nv
        if i in (0, 1, 2, 50, 299):
            for got, want in ((u, ou), (v, ov), (fu, ou), (fv, ov)):
                assert same_bits(got.cpu().numpy(), want, mask), f"step {i}"


def test_wave_step_with_varying_coefficient(dev, corc):
    n, g = 48, 1
    dx = 1.0 / (n - 1)
    rng = np.random.default_rng(5)
    shape = (n + 2, n + 2, 1)
    u0 = rng.uniform(-1, 1, shape + (1,)).astype(np.float32)
    v0 = rng.uniform(-1, 1, shape + (1,)).astype(np.float32)
    c2 = (1 + 0.3 * rng.uniform(-1, 1, shape + (1,))).astype(np.float32)
    ud = to_dev(u0, dev)
    bc = GridBoundary(ud, dx, g)
    bc.dirichlet_BC("ALL", 0.0)
    ws = WaveStep(bc, dx, g, to_dev(c2, dev), 0.3 * dx)
    fu, fv = ws(ud, to_dev(v0, dev))
    fu, fv = ws(fu, fv)
    groups = oracle_groups(bc)
    ou, ov = u0, v0
    w = O.laplacian_weights(dx)
    for i in range(2):
        ou2 = corc.bc_functional(ou, g, groups, dx)
        L = corc.laplacian(ou2, w, np.float32(1.0), g)
        L = (c2 * L).astype(np.float32)
        nv = corc.forward_euler(ov, L, np.float32(0.3 * dx))
        ou = corc.forward_euler(ou2, ov if i == 0 else nv, np.float32(0.3 * dx))
        ov = nv
    mask = ghost_free_mask(shape, g, (1,))
    assert same_bits(fu.cpu().numpy(), ou, mask)
    assert same_bits(fv.cpu().numpy(), ov, mask)
