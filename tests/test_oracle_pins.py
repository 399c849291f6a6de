"""Pins the oracle itself (CPU only).

The reference's own FDM tests assert nothing and it ships no golden vectors for
this path, so each oracle function is checked against what the mathematics
fixes: exactness on polynomials, discrete eigenfunctions, dense brute force vs
scipy Kronecker sums, scipy.ndimage correlation, numpy.gradient, summation by
parts, boundary identities, and whole-loop closed forms for the heat and wave
loops.  The FP64 NumPy oracle is pinned here; the FP32 C restatement is then
pinned to it (agreement within FP32 rounding) and to the FP32 NumPy
restatement bit for bit.  A few tests plant a plausible mistake in a copy of
the oracle and check that the pin catches it.
"""
import numpy as np
import pytest
import scipy.ndimage as ndi
import scipy.sparse as sp

from oracle import np_oracle as O

F64 = np.float64


def coords(n, dx, g):
    return np.arange(-g, n + g, dtype=F64) * dx


def field3(nx, ny, nz, g, fn, dx):
    xs = [coords(n, dx, g) if n > 1 else np.zeros(1) for n in (nx, ny, nz)]
    X, Y, Z = np.meshgrid(*xs, indexing="ij")
    return fn(X, Y, Z)[..., None]


# ------------------------------------------------------------------ Laplacian
@pytest.mark.parametrize("shape", [(7, 9, 1), (6, 5, 8), (11, 1, 1)])
def test_laplacian_exact_on_quadratics(shape):
    dx, g = 0.173, 1
    fn = lambda X, Y, Z: 1.5 * X**2 - 0.75 * Y**2 + 2.25 * Z**2 + 0.3 * X * Y - X + 4
    u = field3(*shape, g, fn, dx)
    lap = O.laplacian(u, O.laplacian_weights(dx, F64), 1.0, g)
    want = 0.0
    for a, c in zip(range(3), (1.5, -0.75, 2.25)):
        if shape[a] > 1:
            want += 2 * c
    inner = lap[O.interior(u.shape[:3], g)]
    np.testing.assert_allclose(inner, want, rtol=1e-9, atol=1e-9)


def test_laplacian_discrete_eigenfunctions():
    n, g = 17, 1
    dx = 1.0 / (n - 1)
    k = (1, 2, 3)
    fn = lambda X, Y, Z: np.sin(k[0] * np.pi * X) * np.sin(k[1] * np.pi * Y) * np.sin(k[2] * np.pi * Z)
    u = field3(n, n, n, g, fn, dx)
    lam = sum((2 * np.cos(kk * np.pi * dx) - 2) / dx**2 for kk in k)
    lap = O.laplacian(u, O.laplacian_weights(dx, F64), 0.5, g)
    inner = O.interior(u.shape[:3], g)
    np.testing.assert_allclose(lap[inner], 0.5 * lam * u[inner], atol=1e-9)


def _dense_operator(shape, g, dx):
    """Matrix of the oracle Laplacian on the interior nodes, ghosts held at 0."""
    P = tuple(s + 2 * g if s > 1 else 1 for s in shape)
    inner = O.interior(P, g)
    n = int(np.prod(shape))
    M = np.zeros((n, n))
    for col in range(n):
        u = np.zeros(P + (1,))
        e = np.zeros(n)
        e[col] = 1.0
        u[inner] = e.reshape(shape + (1,))
        M[:, col] = O.laplacian(u, O.laplacian_weights(dx, F64), 1.0, g)[inner].ravel()
    return M


@pytest.mark.parametrize("shape", [(4, 5, 1), (3, 4, 5)])
def test_laplacian_dense_matches_kronecker_sum(shape):
    dx, g = 0.37, 1
    M = _dense_operator(shape, g, dx)
    dims = [s for s in shape if s > 1]
    T = [sp.diags([1, -2, 1], [-1, 0, 1], shape=(d, d)) / dx**2 for d in dims]
    I = [sp.identity(d) for d in dims]
    K = 0
    for a in range(len(dims)):
        term = None
        for b in range(len(dims)):
            f = T[b] if a == b else I[b]
            term = f if term is None else sp.kron(term, f)
        K = K + term
    np.testing.assert_allclose(M, K.toarray(), rtol=1e-12, atol=1e-9)
    assert np.allclose(M, M.T)                          # symmetric
    assert np.all(np.linalg.eigvalsh(M) < 0)            # negative definite (Dirichlet)


def test_periodic_laplacian_matches_ndimage_wrap():
    rng = np.random.default_rng(3)
    n, g, dx = 9, 1, 0.21
    core = rng.uniform(-1, 1, (n, n, n))
    u = np.zeros((n + 2, n + 2, n + 2, 1))
    u[1:-1, 1:-1, 1:-1, 0] = core
    O.periodic_fill(u, g)
    lap = O.laplacian(u, O.laplacian_weights(dx, F64), 1.0, g)
    k = np.zeros((3, 3, 3))
    k[1, 1, 1] = -6
    for a in range(3):
        idx = [1, 1, 1]
        idx[a] = 0
        k[tuple(idx)] = 1
        idx[a] = 2
        k[tuple(idx)] = 1
    want = ndi.correlate(core, k / dx**2, mode="wrap")
    np.testing.assert_allclose(lap[1:-1, 1:-1, 1:-1, 0], want, rtol=1e-12, atol=1e-9)


def test_laplacian_vector_is_componentwise():
    rng = np.random.default_rng(1)
    u = rng.uniform(-1, 1, (6, 7, 5, 3))
    w = O.laplacian_weights(0.3, F64)
    full = O.laplacian(u, w, 1.0, 1)
    for c in range(3):
        one = O.laplacian(u[..., c:c + 1].copy(), w, 1.0, 1)
        np.testing.assert_array_equal(full[..., c:c + 1], one)


# --------------------------------------------------------------- Grad and Div
def test_grad_exact_on_quadratics_and_numpy_gradient():
    dx, g = 0.11, 1
    fn = lambda X, Y, Z: 0.5 * X**2 + 3 * Y - 2 * Z**2 + X * Z
    u = field3(6, 7, 8, g, fn, dx)
    G = O.grad(u, O.central_weights(dx, F64), 1.0, g)
    xs = [coords(n, dx, g) for n in (6, 7, 8)]
    X, Y, Z = np.meshgrid(*xs, indexing="ij")
    inner = O.interior(u.shape[:3], g)
    exact = [X + Z, 3 + 0 * X, -4 * Z + X]
    for t in range(3):
        np.testing.assert_allclose(G[inner][..., t], exact[t][inner], atol=1e-10)
    ng = np.gradient(u[..., 0], dx, dx, dx)
    for t in range(3):
        np.testing.assert_allclose(G[inner][..., t], ng[t][inner], atol=1e-10)


def test_summation_by_parts_periodic():
    rng = np.random.default_rng(5)
    n, g, dx = 8, 1, 0.4
    P = n + 2 * g
    u = np.zeros((P, P, P, 1))
    v = np.zeros((P, P, P, 3))
    u[1:-1, 1:-1, 1:-1] = rng.uniform(-1, 1, (n, n, n, 1))
    v[1:-1, 1:-1, 1:-1] = rng.uniform(-1, 1, (n, n, n, 3))
    O.periodic_fill(u, g)
    O.periodic_fill(v, g)
    w = O.central_weights(dx, F64)
    gu = O.grad(u, w, 1.0, g)
    dv = O.divergence(v, w, 1.0, g)
    inner = O.interior((P, P, P), g)
    lhs = np.sum(gu[inner] * v[inner])
    rhs = -np.sum(u[inner] * dv[inner])
    assert abs(lhs - rhs) < 1e-10 * max(1.0, abs(lhs))


def test_div_grad_is_wide_laplacian():
    rng = np.random.default_rng(7)
    n, g, dx = 7, 2, 0.3
    P = n + 2 * g
    u = np.zeros((P, P, 1, 1))
    u[g:-g, g:-g, 0, 0] = rng.uniform(-1, 1, (n, n))
    O.periodic_fill(u, g)
    w = O.central_weights(dx, F64)
    gu = O.grad(u, w, 1.0, g)
    O.periodic_fill(gu, g)
    dg = O.divergence(gu, w, 1.0, g)
    core = u[g:-g, g:-g, 0, 0]
    wide = sum(np.roll(core, 2, a) - 2 * core + np.roll(core, -2, a) for a in (0, 1)) / (4 * dx**2)
    np.testing.assert_allclose(dg[g:-g, g:-g, 0, 0], wide, atol=1e-9)


def test_matrix_forms_are_stacked_vector_forms():
    rng = np.random.default_rng(9)
    u = rng.uniform(-1, 1, (6, 5, 1, 2))
    w = O.central_weights(0.2, F64)
    G = O.grad(u, w, 0.7, 1)          # (X, Y, Z, 2, 2)
    for r in range(2):
        np.testing.assert_array_equal(G[..., r, :], O.grad(u[..., r:r + 1].copy(), w, 0.7, 1))
    D = O.divergence(G, w, 1.3, 1, matrix_in=True)
    for r in range(2):
        np.testing.assert_array_equal(D[..., r:r + 1], O.divergence(G[..., r, :].copy(), w, 1.3, 1))


def test_discrete_div_of_discrete_curl_vanishes():
    rng = np.random.default_rng(11)
    n, g, dx = 9, 1, 0.25
    P = n + 2 * g
    psi = np.zeros((P, P, 1, 1))
    psi[1:-1, 1:-1, 0, 0] = rng.uniform(-1, 1, (n, n))
    O.periodic_fill(psi, g)
    w = O.central_weights(dx, F64)
    gp = O.grad(psi, w, 1.0, g)
    vel = np.stack([gp[..., 1], -gp[..., 0]], axis=-1)  # (d psi/dy, -d psi/dx)
    O.periodic_fill(vel, g)
    d = O.divergence(vel, w, 1.0, g)
    assert np.max(np.abs(d[1:-1, 1:-1])) < 1e-12


# ----------------------------------------------------------------- boundaries
def _random_field(shape3, C, seed=0):
    return np.random.default_rng(seed).uniform(-1, 1, shape3 + (C,))


@pytest.mark.parametrize("shape3", [(8, 7, 1), (6, 5, 7)])
def test_legacy_dirichlet_and_neumann_identities(shape3):
    g, dx, C = 1, 0.125, 2
    u = _random_field(shape3, C)
    ijk, groups, inter = O.legacy_boundary_layout(shape3, g)
    bt = np.ones((len(ijk), C), np.int8)
    bv = np.zeros((len(ijk), C))
    bv[:, 0] = 0.75                      # Dirichlet 0.75 on component 0 everywhere
    bt[:, 1] = 2
    bv[:, 1] = -1.5                      # Neumann d/dx_a = -1.5 on component 1
    out = O.bc_legacy(u, ijk, inter, bt, bv, dx)
    Y, Z = shape3[1], shape3[2]
    sn = np.array([Y * Z, Z, 1])
    fo, fi = out.reshape(-1, C), u.reshape(-1, C)
    node = ijk @ sn
    assert np.allclose(fo[node, 0], 0.75)
    for a in range(3):
        b = inter[:, a] != 0
        s = inter[b, a]
        ghost, adj = node[b] - s * sn[a], node[b] + s * sn[a]
        np.testing.assert_allclose((fo[ghost, 0] + fi[adj, 0]) / 2, 0.75, atol=1e-14)
        lo, hi = s > 0, s < 0
        np.testing.assert_allclose((fi[adj[lo], 1] - fo[ghost[lo], 1]) / (2 * dx), -1.5, atol=1e-12)
        np.testing.assert_allclose((fo[ghost[hi], 1] - fi[adj[hi], 1]) / (2 * dx), -1.5, atol=1e-12)


def test_legacy_defaults_to_dirichlet_zero_and_matches_functional_all():
    shape3, g = (9, 8, 1), 1
    u = _random_field(shape3, 1, seed=2)
    ijk, groups, inter = O.legacy_boundary_layout(shape3, g)
    out_l = O.bc_legacy(u, ijk, inter, np.ones((len(ijk), 1), np.int8), np.zeros((len(ijk), 1)), 0.1)
    fg = O.functional_groups(shape3, g)
    out_f = O.bc_functional(u, g, [dict(ids=fg["ALL"], type=1, comps=[0], const=0.0)], 0.1)
    np.testing.assert_array_equal(out_l, out_f)
    assert set(groups) == set(fg)
    for name in ("-X", "+X", "-Y", "+Y"):
        flat_l = np.sort(ijk[groups[name]] @ np.array([8, 1, 1]))
        np.testing.assert_array_equal(flat_l, np.sort(fg[name]))


def test_functional_lid_overlap_quirk_by_hand():
    """No-slip on ALL, then lid u=1 on +Y component 0: the lid group's Dirichlet
    writes the whole node vector (component 1 taken from the input), undoing the
    no-slip zero of component 1 at the lid nodes (reference A.7)."""
    shape3, g, C = (5, 5, 1), 1, 2
    u = np.arange(25 * C, dtype=F64).reshape(shape3 + (C,)) + 1.0
    fg = O.functional_groups(shape3, g)
    groups = [dict(ids=fg["ALL"], type=1, comps=[0, 1], const=0.0),
              dict(ids=fg["+Y"], type=1, comps=[0], const=1.0)]
    out = O.bc_functional(u, g, groups, 0.25)
    for i in range(1, 4):
        assert out[i, 3, 0, 0] == 1.0                      # lid value
        assert out[i, 3, 0, 1] == u[i, 3, 0, 1]            # input restored, not 0
        assert out[i, 4, 0, 0] == 2 * 1.0 - u[i, 2, 0, 0]  # lid ghost
        assert out[i, 4, 0, 1] == -u[i, 2, 0, 1]           # ghost of comp 1 from group 1
    # the lid group holds the corner node (1, 3) too, so it also rewrites that
    # corner's -X ghost (component 0 only); component 1 keeps group 1's value
    assert out[0, 3, 0, 0] == 2 * 1.0 - u[2, 3, 0, 0]
    assert out[0, 3, 0, 1] == -u[2, 3, 0, 1]
    assert out[0, 2, 0, 0] == -u[2, 2, 0, 0]               # below the lid: group 1 only


def test_ramp_value():
    shape3, g = (6, 6, 1), 1
    u = _random_field(shape3, 1, seed=4).astype(np.float32)
    fg = O.functional_groups(shape3, g)
    grp = [dict(ids=fg["-X"], type=1, comps=[0], const=2.0, tmax=0.5)]
    for t, want in ((0.0, 0.0), (0.25, 1.0), (2.0, 2.0)):
        out = O.bc_functional(u, g, grp, np.float32(0.1), t=t)
        assert np.all(out[1, 1:5, 0, 0] == np.float32(want))


# ------------------------------------------------------------- whole loops
def test_heat_loop_closed_form():
    n, g, alpha = 33, 1, 0.1
    dx = 1.0 / (n - 1)
    dt = dx * dx / (4 * alpha)
    u = field3(n, n, 1, g, lambda X, Y, Z: np.sin(np.pi * X) * np.sin(np.pi * Y), dx)
    shape3 = u.shape[:3]
    ijk, _, inter = O.legacy_boundary_layout(shape3, g)
    bt, bv = np.ones((len(ijk), 1), np.int8), np.zeros((len(ijk), 1))
    bc = lambda x: O.bc_legacy(x, ijk, inter, bt, bv, dx)
    w = O.laplacian_weights(dx, F64)
    u0 = u.copy()
    steps = 200
    for _ in range(steps):
        u = O.heat_step(u, w, alpha, g, dt, bc)
    lam = 2 * (2 * np.cos(np.pi * dx) - 2) / dx**2
    inner = O.interior(shape3, g)
    want = (1 + dt * alpha * lam) ** steps * u0[inner]
    # boundary nodes are reset to 0 by the next BC; compare strictly inside
    np.testing.assert_allclose(u[inner][1:-1, 1:-1], want[1:-1, 1:-1], atol=1e-12)


def test_wave_loop_closed_form():
    n, g, c2 = 33, 1, 1.0
    dx = 1.0 / (n - 1)
    dt = 0.5 * dx
    u = field3(n, n, 1, g, lambda X, Y, Z: np.sin(2 * np.pi * X) * np.sin(np.pi * Y), dx)
    v = np.zeros_like(u)
    shape3 = u.shape[:3]
    ijk, _, inter = O.legacy_boundary_layout(shape3, g)
    bt, bv = np.ones((len(ijk), 1), np.int8), np.zeros((len(ijk), 1))
    bc = lambda x: O.bc_legacy(x, ijk, inter, bt, bv, dx)
    w = O.laplacian_weights(dx, F64)
    u0 = u.copy()
    steps = 150
    for i in range(steps):
        u, v = O.wave_step(u, v, w, c2, g, dt, bc, first=(i == 0))
    # Step 1 is explicit Euler (u1 = u0, v1 = dt c2 L u0); from step 2 on the
    # aliasing makes it leapfrog, u^{n+1} - 2u^n + u^{n-1} = mu u^n (n >= 2),
    # with mu = dt^2 c2 lambda, started from u^1 = u0 and u^2 = (1 + 2 mu) u0.
    lam = ((2 * np.cos(2 * np.pi * dx) - 2) + (2 * np.cos(np.pi * dx) - 2)) / dx**2
    mu = c2 * dt * dt * lam
    theta = np.arccos(1 + 0.5 * mu)
    B = 1.5 * mu / np.sin(theta)
    amp = np.cos((steps - 1) * theta) + B * np.sin((steps - 1) * theta)
    inner = O.interior(shape3, g)
    np.testing.assert_allclose(u[inner][1:-1, 1:-1], amp * u0[inner][1:-1, 1:-1], atol=1e-10)


# ------------------------------------------------- planted mistakes are caught
def test_planted_mistakes_fail_the_pins(monkeypatch):
    dx, g = 0.173, 1
    fn = lambda X, Y, Z: 1.5 * X**2 - 0.75 * Y**2 + 2.25 * Z**2
    u = field3(6, 5, 8, g, fn, dx)
    inner = O.interior(u.shape[:3], g)
    # centre weight applied once overall instead of once per axis
    w = O.laplacian_weights(dx, F64)
    acc = 0
    for a in O.eligible_axes(u.shape[:3]):
        acc = acc + w[0] * u[O.interior(u.shape[:3], g, a, -1)] + w[2] * u[O.interior(u.shape[:3], g, a, 1)]
    acc = acc + w[1] * u[inner]
    assert not np.allclose(acc, 2 * (1.5 - 0.75 + 2.25))
    # swapped Grad weights flip the sign of the exact derivative
    wg = O.central_weights(dx, F64)
    G = O.grad(u, (wg[2], wg[1], wg[0]), 1.0, g)
    assert not np.allclose(G[inner][..., 0], 3.0 * field3(6, 5, 8, g, lambda X, Y, Z: X, dx)[inner][..., 0])
    # Neumann with the outward-normal sign convention breaks the +axis identity
    shape3 = (6, 6, 1)
    uu = _random_field(shape3, 1)
    ijk, _, inter = O.legacy_boundary_layout(shape3, g)
    bt = np.full((len(ijk), 1), 2, np.int8)
    bv = np.full((len(ijk), 1), 0.5)
    bad = O.bc_legacy(uu, ijk, -inter, bt, bv, 0.1)
    good = O.bc_legacy(uu, ijk, inter, bt, bv, 0.1)
    assert not np.allclose(bad, good)
