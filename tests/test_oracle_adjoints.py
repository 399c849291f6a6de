"""Pins the oracle's adjoint (transpose) functions with dot-product tests in
FP64: <A x, y> over the interior = <x, A^T y> over the padded field, to
round-off (CPU only)."""
import numpy as np
import pytest

from oracle import np_oracle as O


def inner_dot(a, b, g):
    sl = O.interior(a.shape[:3], g)
    return float(np.sum(a[sl] * b[sl]))


SHAPES = [(9, 8, 1), (6, 7, 5), (11, 1, 1)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("g", [1, 2])
def test_laplacian_adjoint_dot(shape, g):
    rng = np.random.default_rng(0)
    x = rng.standard_normal(shape + (2,))
    y = rng.standard_normal(shape + (2,))
    w = O.laplacian_weights(0.13, np.float64)
    lhs = inner_dot(O.laplacian(x, w, 0.7, g), y, g)
    rhs = float(np.sum(x * O.laplacian_adjoint(y, w, 0.7, g)))
    assert abs(lhs - rhs) <= 1e-11 * max(1.0, abs(lhs))


@pytest.mark.parametrize("shape", SHAPES)
def test_grad_and_div_adjoint_dot(shape):
    rng = np.random.default_rng(1)
    g = 1
    D = len(O.eligible_axes(shape))
    w = O.central_weights(0.2, np.float64)
    for N in (1, 2):
        x = rng.standard_normal(shape + (N,))
        y = rng.standard_normal(shape + (N, D))
        Gx = O.grad(x, w, 1.3, g, force_matrix=True)
        lhs = inner_dot(Gx, y, g)
        rhs = float(np.sum(x * O.grad_adjoint(y, w, 1.3, g)))
        assert abs(lhs - rhs) <= 1e-11 * max(1.0, abs(lhs))
        m = rng.standard_normal(shape + (N, D))
        z = rng.standard_normal(shape + (N,))
        lhs = inner_dot(O.divergence(m, w, -0.4, g, matrix_in=True), z, g)
        rhs = float(np.sum(m * O.divergence_adjoint(z, w, -0.4, g, D)))
        assert abs(lhs - rhs) <= 1e-11 * max(1.0, abs(lhs))


def test_grad_transpose_is_minus_div_inside():
    """Away from the boundary, Grad^T = -Div (central weights are odd)."""
    rng = np.random.default_rng(2)
    shape, g = (12, 11, 1), 1
    w = O.central_weights(0.1, np.float64)
    y = rng.standard_normal(shape + (1, 2))
    gt = O.grad_adjoint(y, w, 1.0, g)
    dv = O.divergence(y[..., 0, :].copy(), w, -1.0, g)
    core = (slice(2, -2), slice(2, -2))
    np.testing.assert_allclose(gt[core], dv[core], atol=1e-12)


def test_laplacian_self_adjoint_inside():
    rng = np.random.default_rng(3)
    shape, g = (10, 9, 8), 1
    w = O.laplacian_weights(0.1, np.float64)
    y = rng.standard_normal(shape + (1,))
    core = (slice(2, -2),) * 3
    np.testing.assert_allclose(O.laplacian_adjoint(y, w, 1.0, g)[core],
                               O.laplacian(y, w, 1.0, g)[core], atol=1e-9)


def test_boundary_linear_adjoint_dot():
    rng = np.random.default_rng(4)
    shape3, C = (8, 7, 1), 2
    fg = O.functional_groups(shape3, 1)
    groups = [dict(ids=fg["ALL"], type=1, comps=[0, 1], const=0.0),
              dict(ids=fg["-X"], type=2, comps=[1], const=0.0)]
    x = rng.standard_normal(shape3 + (C,))
    y = rng.standard_normal(shape3 + (C,))
    # linear part: constants are 0 here, so bc_functional is linear in x
    Bx = O.bc_functional(x, 1, groups, 0.1)
    # recover the (dst, src, sign) list by probing unit vectors
    n = x.size
    M = np.zeros((n, n))
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1
        M[:, i] = O.bc_functional(e.reshape(x.shape), 1, groups, 0.1).ravel()
    writes = []
    for d in range(n):
        row = M[d]
        nz = np.nonzero(row)[0]
        if len(nz) == 1 and nz[0] == d and row[d] == 1:
            continue
        if len(nz) == 0:
            writes.append((d, 0, 0))
        else:
            assert len(nz) == 1
            writes.append((d, int(nz[0]), float(row[nz[0]])))
    lhs = float(np.sum(Bx * y))
    rhs = float(np.sum(x * O.boundary_linear_adjoint(y, writes)))
    assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs))


def test_heat_step_adjoint_dot():
    rng = np.random.default_rng(5)
    shape, g = (9, 10, 1), 1
    w = O.laplacian_weights(0.1, np.float64)
    x = rng.standard_normal(shape + (1,))
    y = rng.standard_normal(shape + (1,))
    lap = O.laplacian(x, w, 0.3, g)
    lap = np.where(np.isnan(lap), 0.0, lap)   # ghosts: u_next = u2 there
    fx = x + 0.01 * lap
    lhs = float(np.sum(fx * y))
    rhs = float(np.sum(x * O.heat_step_adjoint(y, w, 0.3, g, 0.01)))
    assert abs(lhs - rhs) <= 1e-11 * max(1.0, abs(lhs))


@pytest.mark.parametrize("first", [True, False])
@pytest.mark.parametrize("with_c2", [True, False])
def test_wave_step_adjoint_dot_and_c2_derivative(first, with_c2):
    rng = np.random.default_rng(6)
    shape, g, dt_step, alpha = (10, 9, 1), 1, 0.05, 0.8
    w = O.laplacian_weights(0.1, np.float64)
    u2, v = rng.standard_normal(shape + (1,)), rng.standard_normal(shape + (1,))
    c2 = 1 + 0.3 * rng.standard_normal(shape + (1,)) if with_c2 else None
    gun, gvn = rng.standard_normal(shape + (1,)), rng.standard_normal(shape + (1,))
    un, vn = O.wave_step_core(u2, v, c2, w, alpha, g, dt_step, first)
    gu2, gv, dc2 = O.wave_step_adjoint(gun, gvn, c2, u2, w, alpha, g, dt_step, first)
    lhs = float(np.sum(un * gun) + np.sum(vn * gvn))
    rhs = float(np.sum(u2 * gu2) + np.sum(v * gv))
    assert abs(lhs - rhs) <= 1e-11 * max(1.0, abs(lhs))
    if with_c2:
        # the pairing is affine in c2, so a central difference is exact up to round-off
        dc = rng.standard_normal(shape + (1,))
        eps = 1e-3
        up = O.wave_step_core(u2, v, c2 + eps * dc, w, alpha, g, dt_step, first)
        um = O.wave_step_core(u2, v, c2 - eps * dc, w, alpha, g, dt_step, first)
        fd = (np.sum(up[0] * gun) + np.sum(up[1] * gvn) - np.sum(um[0] * gun) - np.sum(um[1] * gvn)) / (2 * eps)
        assert abs(fd - np.sum(dc2 * dc)) <= 1e-8 * max(1.0, abs(fd))
