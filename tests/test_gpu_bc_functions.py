"""GPU checks of boundary values given by functions: the sine record
(evaluated in the boundary kernels from the device time) and Python boundary
functions (evaluated with torch on the device into the time array the kernels
read), in the functional and legacy GridBoundary, against the oracle's
per-node restatement of the reference's function kernels
(pm/FDM/boundary/functional_boundary.py:356-400,
pm/deprecated/func_boundary.py:96-134).

Tolerance: every scalar the boundary does not write, and every value from a
constant or ramp record, is bit-exact.  A sine value comes from the device
sinf (or torch's device sin) where the oracle rounds the FP64 sine once, so
written scalars are compared within 1e-6 of the values' scale (the per-step
tolerance the path is held to) and must differ by at most a few ulps."""
import numpy as np
import pytest
import torch

from oracle import np_oracle as O
from pde_module_amd import ForwardEuler, HeatStep, Laplacian
from pde_module_amd.deprecated import GridBoundary as LegacyGridBoundary
from pde_module_amd.FDM import GridBoundary
from pde_module_amd.FDM.boundary import get_ramp_func, get_sine_func

from test_bc_functions import _coords, coord_profile, pulse

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    return torch.equal(a.contiguous().view(torch.int32), b.contiguous().view(torch.int32))


def close(got, want, scale):
    got = np.asarray(got, np.float32)
    want = np.asarray(want, np.float32)
    same = got.view(np.uint32) == want.view(np.uint32)
    return bool(np.all(same | (np.abs(got.astype(np.float64) - want) <= 1e-6 * scale)))


@pytest.mark.parametrize("C", [1, 3])
def test_functional_boundary_with_sine_and_functions(dev, C):
    print(f"functional boundary, sine + functions, C={C}", flush=True)
    shape3, g, dx = (34, 21, 1), 1, 0.05
    rng = np.random.default_rng(5)
    u = rng.uniform(-1, 1, shape3 + (C,)).astype(np.float32)
    crd = _coords(shape3, np.float32(dx))
    bc = GridBoundary(torch.from_numpy(u).to(dev), dx, g, grid_coordinates=crd)
    bc.dirichlet_BC("ALL", 0.25)
    bc.dirichlet_BC("-X", get_sine_func(1.5, 4.0), 0)
    cf = (lambda *a: coord_profile(*a)[..., :C]) if C == 3 else \
        (lambda cv, n, co, t, d, p: coord_profile(cv, n, co, t, d, p)[..., :1] * p)
    bc.vonNeumann_BC("+Y", cf, 0)
    bc.dirichlet_BC("+X", get_ramp_func(2.0, 1.0))
    names = list(bc.func_groups)
    grp = O.functional_groups(shape3, g)
    og = [dict(ids=grp["ALL"], type=1, comps=list(range(C)), const=0.25),
          dict(ids=grp["-X"], type=1, comps=[0], const=0.0, sine=(1.5, 4.0)),
          dict(ids=grp["+Y"], type=2, comps=[0], func=cf, name=names[2]),
          dict(ids=grp["+X"], type=1, comps=list(range(C)), const=2.0, tmax=1.0)]
    ud = torch.from_numpy(u).to(dev)
    for t in (0.0, 0.37, 1.9):
        params = {names[2]: 2.0}
        got = bc(ud, t, dict(params)).cpu().numpy()
        want = O.bc_functional(u, g, og, np.float32(dx), t=t, coords=crd, params=params)
        assert close(got, want, 4.0), t
    # a device time array advanced in place: the same values as the float time
    tt = torch.full((1,), 0.37, dtype=torch.float32, device=dev)
    a = bc(ud, tt, {names[2]: 2.0}).clone()
    b = bc(ud, 0.37, {names[2]: 2.0})
    assert bits_equal(a, b)


def test_legacy_interference_loop_matches_oracle_and_graph_replay(dev):
    """The interference example's loop without its slot and damping layer
    (examples/wave2d/Interference.py:113-144): Dirichlet 0 on the walls, the
    -X wall a sine pulse (a Python function reading t and params, then the
    node's own value after t = 2), shifted into the field, leapfrog-like
    u/v update with ForwardEuler; against the oracle loop, then the same loop
    captured as a HIP graph with t in device memory, replayed bit-identical
    to the eager loop."""
    print("legacy interference-style loop with a function boundary", flush=True)
    n, m, g = 41, 2, 1
    dx = 2.0 / (n - 1)
    shape3 = (n * m + 2 * g, n + 2 * g, 1)
    c2, dt, omega, steps = 0.01, 0.3 / (0.1 * np.sqrt(2.0 / dx ** 2)), 40 * np.pi / 4 * 0.1, 60
    dt = float(np.float32(dt))
    u0 = np.zeros(shape3 + (1,), np.float32)
    dev_u = torch.from_numpy(u0).to(dev)

    def make_bc():
        bc = LegacyGridBoundary(dev_u, dx, g)
        bc.dirichlet_BC("ALL", 0.0)
        bc.dirichlet_BC("-X", pulse, 0)
        bc.shift_group("-X", 4, 0)
        return bc

    bc = make_bc()
    lap = Laplacian(1, dx, g)
    # the Laplacian's output ghosts are never written (unzeroed, as the
    # reference's wp.empty outputs); the shifted -X group leaves the x = 0
    # ghost column to the loop, so zero them once as the oracle's 0 ghosts
    lap(dev_u, c2)
    lap.output_array.zero_()
    us, vs = ForwardEuler(), ForwardEuler()
    u, v = dev_u.clone(), dev_u.clone()
    t = 0.0
    for _ in range(steps):
        u3 = bc(u, t, {"-X": omega})
        st = lap(u3, c2)
        vn = vs(v, st, dt)
        un = us(u3, v, dt)
        u, v = un.clone(), vn.clone()
        t = float(np.float32(t) + np.float32(dt))
    got_u = u.cpu().numpy()
    # oracle loop (FP32 NumPy, per-node boundary function)
    ou, ov = u0.copy(), u0.copy()
    w = O.laplacian_weights(np.float32(dx))
    ot = 0.0
    for _ in range(steps):
        u3 = O.bc_legacy(ou, bc.boundary_ijk_indices, bc.boundary_interior, bc.boundary_type,
                         bc.boundary_value, np.float32(dx), t=ot, params={"-X": omega},
                         func_groups=[dict(ids=bc.groups["-X"], type=1, var=0, func=pulse,
                                           name="-X")])
        st = O.laplacian(u3, w, np.float32(c2), g)
        st = np.where(np.isnan(st), np.float32(0), st)
        ov, ou = O.forward_euler(ov, st, np.float32(dt)), O.forward_euler(u3, ov, np.float32(dt))
        ot = float(np.float32(ot) + np.float32(dt))
    inner = (slice(g, -g), slice(g, -g))
    scale = float(np.abs(ou[inner]).max())
    assert scale > 0.1
    err = float(np.abs(got_u[inner].astype(np.float64) - ou[inner]).max())
    print(f"  loop vs oracle after {steps} steps: max abs {err:.2e} (scale {scale:.2f})",
          flush=True)
    assert err <= 1e-5 * scale

    # graph-captured loop: t in device memory, advanced by the graph itself
    bcg = make_bc()
    lapg = Laplacian(1, dx, g)
    ug, vg = dev_u.clone(), dev_u.clone()
    tt = torch.zeros(1, dtype=torch.float32, device=dev)
    dtt = torch.full((1,), dt, dtype=torch.float32, device=dev)
    un_buf, vn_buf = torch.empty_like(ug), torch.empty_like(vg)
    usg, vsg = ForwardEuler(), ForwardEuler()

    def one_step():
        u3 = bcg(ug, tt, {"-X": omega})
        st = lapg(u3, c2)
        vn_buf.copy_(vsg(vg, st, dt))
        un_buf.copy_(usg(u3, vg, dt))
        ug.copy_(un_buf)
        vg.copy_(vn_buf)
        tt.add_(dtt)

    one_step()                     # set-up (tables, outputs) outside the capture
    lapg.output_array.zero_()
    torch.cuda.synchronize()
    ug.copy_(dev_u)
    vg.copy_(dev_u)
    tt.zero_()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        one_step()
    ug.copy_(dev_u)
    vg.copy_(dev_u)
    tt.zero_()
    torch.cuda.synchronize()
    for _ in range(steps):
        graph.replay()
    torch.cuda.synchronize()
    assert bits_equal(ug, u), "graph-replayed loop differs from the eager loop"


def test_one_launch_heat_step_with_function_values(dev):
    """HeatStep's one-launch path (boundary evaluated in the ring pieces)
    reads sine and array values from the same time array as the in-place
    boundary + step: identical bits for every state."""
    print("one-launch heat step with sine and function boundary values", flush=True)
    shape, g = (70, 90), 1
    rng = np.random.default_rng(9)
    u0 = torch.from_numpy(rng.uniform(-1, 1, shape + (1, 1)).astype(np.float32)).to(dev)
    dx, alpha = 0.01, 0.05
    dt = 0.2 * dx * dx / alpha
    crd = _coords(shape + (1,), np.float32(dx))

    def make(u):
        bc = GridBoundary(u, dx, g, grid_coordinates=crd)
        bc.dirichlet_BC("ALL", 0.0)
        bc.dirichlet_BC("-Y", get_sine_func(2.0, 30.0))
        bc.vonNeumann_BC("+X", lambda cv, nid, co, t, d, p: co[nid[:, 0], nid[:, 1], 0, 1:2]
                         * torch.cos(5.0 * t))
        return bc

    one = HeatStep(make(u0), dx, g, alpha, dt)
    two = HeatStep(make(u0), dx, g, alpha, dt, one_launch=False)
    ua, ub = u0.clone(), u0.clone()
    for k in range(6):
        t = 0.05 * k
        na, nb = one(ua, t), two(ub, t)
        torch.cuda.synchronize()
        assert one._ring[1] is not None, "the one-launch path was not taken"
        assert bits_equal(na, nb), f"step {k}"
        ua, ub = na, nb
