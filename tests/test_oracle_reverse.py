"""Pins, on the CPU, the time reversal behind the wave loop's reversed
gradient (``WaveRollout(reverse=True)``, ``pde_wave_reverse_adjoint_pair_f32``):
with the NumPy FP64 oracle's own step (``wave_step_core``, the reference loop
``examples/wave2d/standingWave.py:70-76``) and legacy boundary
(``bc_legacy``, ``pm/deprecated/gridBoundary.py:267-302``), walking the loop
backwards -- u2 = u_next - dt*v_next and v = v_next - dt*c2*alpha*lap(u2) -
source on interior nodes, ghosts copied, then the scalars the boundary
overwrote put back -- rebuilds every forward state but the first to FP64
rounding.  Without the recorded scalars it does not, which is why the forward
records them."""
import numpy as np

from oracle import np_oracle as O


def _setup(seed=3, steps=40):
    rng = np.random.default_rng(seed)
    shape, g, dx, alpha = (13, 11, 1), 1, 0.1, 0.9
    ijk, _, inter = O.legacy_boundary_layout(shape, g)
    nb = len(ijk)
    bt = np.ones((nb, 1), np.int8)
    bt[rng.random(nb) < 0.4] = 2                 # some Neumann entries
    bv = rng.standard_normal((nb, 1))
    bc = lambda x: O.bc_legacy(x, ijk, inter, bt, bv, dx)
    u0 = rng.standard_normal(shape + (1,))
    v0 = rng.standard_normal(shape + (1,))
    c2 = 1 + 0.3 * rng.uniform(-1, 1, shape + (1,))
    w = O.laplacian_weights(dx, np.float64)
    dt = 0.3 * dx
    # a source inside, one on a boundary node, one repeated
    src = np.array([6 * 11 + 5, 1 * 11 + 4, 6 * 11 + 5], np.int64)
    wav = rng.standard_normal((steps, len(src)))
    # scalars the boundary writes: where it changes a random field
    r = rng.standard_normal(shape + (1,))
    W = bc(r) != r
    return shape, g, alpha, bc, u0, v0, c2, w, dt, src, wav, W


def _forward(steps, g, alpha, bc, u0, v0, c2, w, dt, src, wav, W):
    u, v = u0, v0
    us, vs, saved = [u0], [v0], []
    for n in range(steps):
        saved.append(u[W].copy())                # the forward's record
        u2 = bc(u)
        v = v.copy()
        np.add.at(v.reshape(-1), src, dt * wav[n])
        u, v = O.wave_step_core(u2, v, c2, w, alpha, g, dt, n == 0)
        us.append(u)
        vs.append(v)
    return us, vs, saved


def _back(u, v, n, shape, g, alpha, c2, w, dt, src, wav):
    """Step n backwards from (u_{n+1}, v_{n+1}): (u2_n, v_n)."""
    M = O.interior_mask(shape, g)[..., None]
    u2 = np.where(M, u - dt * v, u)
    vp = np.where(M, v - dt * (c2 * O.laplacian(u2, w, alpha, g)), v)
    np.subtract.at(vp.reshape(-1), src[::-1], dt * wav[n][::-1])
    return u2, vp


def test_reverse_walk_rebuilds_every_state_but_the_first():
    steps = 40
    shape, g, alpha, bc, u0, v0, c2, w, dt, src, wav, W = _setup(steps=steps)
    us, vs, saved = _forward(steps, g, alpha, bc, u0, v0, c2, w, dt, src, wav, W)
    u, v = us[-1], vs[-1]
    scale = max(np.abs(np.stack(us)).max(), np.abs(np.stack(vs)).max())
    for n in range(steps - 1, 0, -1):
        u2, vp = _back(u, v, n, shape, g, alpha, c2, w, dt, src, wav)
        # u2 is the boundary-applied state step n read
        np.testing.assert_allclose(u2, bc(us[n]), rtol=0, atol=1e-10 * scale)
        up = u2.copy()
        up[W] = saved[n]
        np.testing.assert_allclose(up, us[n], rtol=0, atol=1e-10 * scale)
        np.testing.assert_allclose(vp, vs[n], rtol=0, atol=1e-10 * scale)
        u, v = up, vp


def test_reverse_walk_needs_the_recorded_scalars():
    steps = 12
    shape, g, alpha, bc, u0, v0, c2, w, dt, src, wav, W = _setup(seed=5, steps=steps)
    us, vs, saved = _forward(steps, g, alpha, bc, u0, v0, c2, w, dt, src, wav, W)
    u, v = us[-1], vs[-1]
    for n in range(steps - 1, steps - 4, -1):
        u, v = _back(u, v, n, shape, g, alpha, c2, w, dt, src, wav)   # boundary left applied
    assert np.abs(u - us[steps - 3]).max() > 1e-3
