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


# ----------------------------------------------------------------------------
# Independent pin of the composed step.  A quadratic velocity field makes the
# central differences and the compact Laplacian exact, so the predictor has a
# closed form; the pressure is a scipy sparse direct solve of the compact
# Laplacian with p = 0 on the ghost layer; the correction is numpy.gradient of
# that pressure.  None of these reuses the oracle's operators.
_CX = (0.3, 0.5, -0.2, 0.4, -0.3, 0.2)      # u_x = c0 + c1 x + c2 y + c3 x^2 + c4 xy + c5 y^2
_CY = (-0.1, 0.2, 0.6, -0.3, 0.5, 0.1)


def _quad(c, x, y):
    return c[0] + c[1] * x + c[2] * y + c[3] * x * x + c[4] * x * y + c[5] * y * y


def _quad_dx(c, x, y):
    return c[1] + 2 * c[3] * x + c[4] * y


def _quad_dy(c, x, y):
    return c[2] + c[4] * x + 2 * c[5] * y


def _quad_lap(c):
    return 2 * c[3] + 2 * c[5]


def _cubic_extrapolated_ghosts(a, g=1):
    """Ghost layer g=1 by cubic extrapolation along each axis (exact for the
    cubic fields this pin produces)."""
    b = np.array(a, copy=True)
    P0, P1 = b.shape[:2]
    inner1 = slice(1, P1 - 1)
    b[0, inner1] = 4 * b[1, inner1] - 6 * b[2, inner1] + 4 * b[3, inner1] - b[4, inner1]
    b[-1, inner1] = 4 * b[-2, inner1] - 6 * b[-3, inner1] + 4 * b[-4, inner1] - b[-5, inner1]
    b[:, 0] = 4 * b[:, 1] - 6 * b[:, 2] + 4 * b[:, 3] - b[:, 4]
    b[:, -1] = 4 * b[:, -2] - 6 * b[:, -3] + 4 * b[:, -4] - b[:, -5]
    return b


def _zero_ghosts(a):
    b = np.zeros_like(a)
    b[1:-1, 1:-1] = a[1:-1, 1:-1]
    return b


def _closed_form_step(n, nu, dt):
    """u^{n+1} and p for one projection step of the quadratic field, from the
    closed-form predictor, a sparse direct Poisson solve and numpy.gradient."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    P = n + 2
    dx = 1.0 / (n - 1)
    x = (np.arange(P) - 1) * dx
    X, Y = np.meshgrid(x, x, indexing="ij")
    ux, uy = _quad(_CX, X, Y), _quad(_CY, X, Y)
    conv_x = ux * _quad_dx(_CX, X, Y) + uy * _quad_dy(_CX, X, Y)
    conv_y = ux * _quad_dx(_CY, X, Y) + uy * _quad_dy(_CY, X, Y)
    usx = ux + dt * (-conv_x + nu * _quad_lap(_CX))
    usy = uy + dt * (-conv_y + nu * _quad_lap(_CY))
    rhs = (np.gradient(usx, dx, axis=0) + np.gradient(usy, dx, axis=1)) / dt
    T = sp.diags([np.ones(n - 1), -2 * np.ones(n), np.ones(n - 1)], [-1, 0, 1])
    I = sp.identity(n)
    L = (sp.kron(T, I) + sp.kron(I, T)).tocsc() / (dx * dx)
    p = np.zeros((P, P))
    p[1:-1, 1:-1] = spla.spsolve(L, rhs[1:-1, 1:-1].reshape(-1)).reshape(n, n)
    unx = usx - dt * np.gradient(p, dx, axis=0)
    uny = usy - dt * np.gradient(p, dx, axis=1)
    u0 = np.stack([ux, uy], -1).reshape(P, P, 1, 2)
    return u0, np.stack([unx, uny], -1).reshape(P, P, 1, 2), p.reshape(P, P, 1, 1), rhs


def _composition(u, p, lw, cw, nu, dt_step, h2, inv, sweeps, g, ubc, pbc, mistake=None):
    """The oracle's step built from its primitives, with one planted mistake:
    the checker below must reject every variant but mistake=None."""
    d = u.dtype.type
    u2 = ubc(u)
    du = O.grad(u2, cw, d(1.0), g)
    if mistake == "transposed_matvec":
        du = np.swapaxes(du, -1, -2)
    conv = O.matvec(du, u2)
    diff = O.laplacian(u2, lw, d(nu), g)
    F = O.linear_combination((+1 if mistake == "plus_conv" else -1, conv), (1, diff))
    us = O.forward_euler(u2, F, dt_step)
    us2 = ubc(us)
    rhs = O.divergence(us2, cw, d(1.0 if mistake == "no_inv_dt" else 1.0 / dt_step), g)
    for _ in range(sweeps):
        p = O.jacobi(pbc(p), rhs, g, h2, inv)
    gp = O.grad(pbc(p), cw, d(1.0), g)
    return O.forward_euler(us2, gp, dt_step if mistake == "plus_correction" else -dt_step), p


_N, _NU, _DT = 16, 0.05, 0.01


def _check_against_closed_form(step):
    n = _N
    dx = 1.0 / (n - 1)
    u0, want_u, want_p, _ = _closed_form_step(n, _NU, _DT)
    lw, cw = O.laplacian_weights(dx, np.float64), O.central_weights(dx, np.float64)
    p0 = np.zeros_like(want_p)
    # 3000 Jacobi sweeps: spectral radius cos(pi/17) ~ 0.983, so ~1e-22 of the start error
    got_u, got_p = step(u0, p0, lw, cw, _NU, _DT, dx * dx, 0.25, 3000, 1,
                        _cubic_extrapolated_ghosts, _zero_ghosts)
    inner = (slice(1, -1), slice(1, -1))
    scale_u = np.abs(want_u[inner]).max()
    scale_p = np.abs(want_p[inner]).max()
    return (np.abs(got_u[inner] - want_u[inner]).max() / scale_u,
            np.abs(got_p[inner] - want_p[inner]).max() / scale_p)


def test_projection_step_matches_closed_form_and_direct_poisson_solve():
    eu, ep = _check_against_closed_form(O.projection_step)
    assert eu < 1e-10 and ep < 1e-10, (eu, ep)


@pytest.mark.parametrize("mistake", ["plus_correction", "plus_conv", "no_inv_dt",
                                     "transposed_matvec"])
def test_planted_projection_mistakes_fail_the_closed_form(mistake):
    ok_u, ok_p = _check_against_closed_form(
        lambda *a: _composition(*a, mistake=None))
    assert ok_u < 1e-10 and ok_p < 1e-10
    eu, ep = _check_against_closed_form(lambda *a: _composition(*a, mistake=mistake))
    assert max(eu, ep) > 1e-3, (mistake, eu, ep)


@pytest.mark.parametrize("mistake", [None, "plus_correction", "no_inv_dt"])
def test_projection_removes_most_of_the_divergence(mistake):
    """Physical check: with a converged pressure the corrected field keeps only
    the compact-vs-wide Laplacian mismatch of the divergence (here < 5 % of the
    predictor's); a flipped correction roughly doubles it, a missing 1/dt
    leaves it almost unchanged."""
    n = _N
    dx = 1.0 / (n - 1)
    u0, _, _, rhs = _closed_form_step(n, _NU, _DT)
    lw, cw = O.laplacian_weights(dx, np.float64), O.central_weights(dx, np.float64)
    un, _ = _composition(u0, np.zeros(u0.shape[:3] + (1,)), lw, cw, _NU, _DT, dx * dx, 0.25,
                         3000, 1, _cubic_extrapolated_ghosts, _zero_ghosts, mistake=mistake)
    inner2 = (slice(2, -2), slice(2, -2))
    d_before = np.abs(rhs[inner2] * _DT).max()
    d_after = np.abs((np.gradient(un[..., 0, 0], dx, axis=0)
                      + np.gradient(un[..., 0, 1], dx, axis=1))[inner2]).max()
    ratio = d_after / d_before
    if mistake is None:
        assert ratio < 0.05, ratio
    else:
        assert ratio > 0.5, (mistake, ratio)
