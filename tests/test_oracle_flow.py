"""Pins the projection cavity's oracle pieces (CPU only): the Jacobi sweep
keeps a discrete solution of lap(p) = rhs fixed (quadratic, exactly in FP64),
its transpose passes the dot-product test, and the composed step keeps the
zero state at rest without a lid."""
import numpy as np
import pytest

from oracle import np_oracle as O
from pde_module_amd.workloads import ns_cavity_u0


def _jacobi_adjoint(gout, g, h2, inv):
    """Transposed Jacobi written from the definition (dense loop)."""
    shape3 = gout.shape[:3]
    P = gout.shape
    gp = np.zeros_like(gout)
    grhs = np.zeros_like(gout)
    inner = O.interior(shape3, g)
    gi = np.zeros_like(gout)
    gi[inner] = gout[inner]
    for a in O.eligible_axes(shape3):
        for s in (-1, 1):
            # out[n] reads p[n + s e]: scatter inv * g[n] to n + s e
            src = O.interior(shape3, g)
            dst = O.interior(shape3, g, a, s)
            gp[dst] += inv * gi[src]
    grhs[inner] = -h2 * inv * gi[inner]
    return gp, grhs


@pytest.mark.parametrize("shape", [(9, 8, 1), (6, 7, 5), (10, 1, 1)])
def test_jacobi_fixed_point_on_quadratics(shape):
    g = 1
    h = 0.1
    X, Y, Z = np.meshgrid(*[np.arange(s) * h for s in shape], indexing="ij")
    axes = O.eligible_axes(shape)
    p = sum([X, Y, Z][a] ** 2 for a in axes)[..., None].astype(np.float64)
    rhs = np.full_like(p, 2.0 * len(axes))
    inv = 1.0 / (2 * len(axes))
    out = O.jacobi(p, rhs, g, h * h, inv)
    inner = O.interior(shape, g)
    assert np.allclose(out[inner], p[inner], rtol=0, atol=1e-12)
    assert np.isnan(out[~O.interior_mask(shape, g)]).all()


@pytest.mark.parametrize("shape", [(9, 8, 1), (6, 7, 5)])
def test_jacobi_transpose_dot(shape):
    rng = np.random.default_rng(1)
    g, h2 = 1, 0.01
    inv = 1.0 / (2 * len(O.eligible_axes(shape)))
    p = rng.standard_normal(shape + (1,))
    rhs = rng.standard_normal(shape + (1,))
    y = rng.standard_normal(shape + (1,))
    out = O.jacobi(p, rhs, g, h2, inv)
    inner = O.interior_mask(shape, g)
    lhs = np.sum(np.where(inner[..., None], out, 0.0) * y)
    gp, grhs = _jacobi_adjoint(y, g, h2, inv)
    assert np.isclose(lhs, np.sum(gp * p) + np.sum(grhs * rhs), rtol=1e-12)


def test_projection_step_rest_state_and_divergence_free_start():
    n, g = 12, 1
    P = n + 2 * g
    dx = 1.0 / (n - 1)
    lw, cw = O.laplacian_weights(dx, np.float64), O.central_weights(dx, np.float64)
    m = O.interior_mask((P, P, 1), g)
    zero_ghosts = lambda a: np.where(m.reshape(m.shape + (1,) * (a.ndim - 3)), a, 0.0)
    u0 = np.zeros((P, P, 1, 2))
    p0 = np.zeros((P, P, 1, 1))
    u1, p1 = O.projection_step(u0, p0, lw, cw, 0.01, 1e-3, dx * dx, 0.25, 5, g, zero_ghosts,
                               zero_ghosts)
    inner = O.interior((P, P, 1), g)
    assert np.all(u1[inner] == 0) and np.all(p1[inner] == 0)
    u = ns_cavity_u0(3, n)
    assert u.shape == (P, P, 1, 2) and np.isclose(np.abs(u).max(), 1e-3)
    d = O.divergence(u.astype(np.float64), cw, 1.0, g)
    # FP32 rounding of a 1e-3 field only: 1e-6 of the 1e-3 / dx scale of its derivatives
    assert np.abs(d[O.interior((P, P, 1), 2)]).max() < 1e-6 * 1e-3 / dx
