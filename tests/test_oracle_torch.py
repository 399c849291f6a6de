"""Pins the FP64 torch restatement of the loops (used for reference
gradients) against the NumPy FP64 oracle, and its autograd against central
finite differences (CPU only)."""
import numpy as np
import torch

from oracle import np_oracle as O
from oracle import torch_fp64 as TF


def _setup(n=20, seed=0):
    rng = np.random.default_rng(seed)
    shape = (n + 2, n + 2, 1)
    u0 = rng.uniform(-1, 1, shape + (1,))
    v0 = rng.uniform(-1, 1, shape + (1,))
    c2 = 1 + 0.3 * rng.uniform(-1, 1, shape + (1,))
    return shape, u0, v0, c2


def test_heat_and_wave_match_numpy_oracle():
    shape, u0, v0, c2 = _setup()
    dx, g = 1.0 / 19, 1
    ijk, _, inter = O.legacy_boundary_layout(shape, g)
    bt, bv = np.ones((len(ijk), 1), np.int8), np.zeros((len(ijk), 1))
    bc = lambda x: O.bc_legacy(x, ijk, inter, bt, bv, dx)
    w = O.laplacian_weights(dx, np.float64)
    u = u0.copy()
    for _ in range(15):
        u = O.heat_step(u, w, 0.1, g, 0.2 * dx * dx / 0.1, bc)
    ut = TF.heat_rollout(torch.from_numpy(u0), dx, g, 0.1, 0.2 * dx * dx / 0.1, 15).numpy()
    m = O.interior_mask(shape, g)
    np.testing.assert_allclose(ut[m], u[m], rtol=1e-12, atol=1e-12)
    uu, vv = u0.copy(), v0.copy()
    dt = 0.3 * dx
    for i in range(12):
        u2 = bc(uu)
        uu, vv = O.wave_step_core(u2, vv, c2, w, 1.0, g, dt, i == 0)
    ut, vt, _ = TF.wave_rollout(torch.from_numpy(c2), torch.from_numpy(u0), torch.from_numpy(v0),
                                dx, g, dt, 12)
    np.testing.assert_allclose(ut.numpy()[m], uu[m], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(vt.numpy()[m], vv[m], rtol=1e-12, atol=1e-12)


def test_autograd_matches_finite_differences():
    shape, u0, v0, c2 = _setup(12, 1)
    dx, g, dt, steps = 1.0 / 11, 1, 0.3 / 11, 20
    rec_idx = torch.tensor([15, 40, 77], dtype=torch.int64)
    src_idx = torch.tensor([85], dtype=torch.int64)
    wav = torch.linspace(0, 1, steps, dtype=torch.float64)[:, None]
    d = torch.randn(steps, 3, dtype=torch.float64, generator=torch.Generator().manual_seed(0))

    def J(c):
        _, _, rec = TF.wave_rollout(c, torch.from_numpy(u0), torch.from_numpy(v0), dx, g, dt,
                                    steps, src_idx=src_idx, src_wavelet=wav, rec_idx=rec_idx)
        return ((rec - d) ** 2).sum()

    c = torch.from_numpy(c2).clone().requires_grad_(True)
    grad, = torch.autograd.grad(J(c), c)
    dirn = torch.from_numpy(np.random.default_rng(2).standard_normal(c2.shape))
    eps = 1e-6
    with torch.no_grad():
        fd = (J(c + eps * dirn) - J(c - eps * dirn)) / (2 * eps)
    assert abs(float((grad * dirn).sum()) - float(fd)) <= 1e-7 * max(1.0, abs(float(fd)))
