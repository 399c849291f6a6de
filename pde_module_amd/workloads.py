"""Seeded synthetic inputs of the benchmark workloads (NumPy
``default_rng(seed)``, generated in FP64 and rounded to FP32).

* heat_2d_config:   256 x 256 nodes (+1 ghost ring), dx = 1/255, alpha = 0.01,
                    dt = 0.2 dx^2 / alpha, 8 Gaussian bumps, Dirichlet 0 walls.
* operator_3d_fields: periodic scalar and vec3 fields, i.i.d. U[-1, 1].
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def gaussian_bumps(rng, coords_x, coords_y, n_bumps=8):
    """sum_b A_b exp(-|x - c_b|^2 / (2 w_b^2)) with centres U[0.2, 0.8]^2,
    widths U[0.02, 0.08], amplitudes U[0.5, 1], drawn in that order."""
    centres = rng.uniform(0.2, 0.8, size=(n_bumps, 2))
    widths = rng.uniform(0.02, 0.08, size=n_bumps)
    amps = rng.uniform(0.5, 1.0, size=n_bumps)
    X, Y = np.meshgrid(coords_x, coords_y, indexing="ij")
    u = np.zeros_like(X)
    for (cx, cy), w, a in zip(centres, widths, amps):
        u += a * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2.0 * w * w))
    return u


@dataclass
class Heat2D:
    n: int
    g: int
    dx: float
    alpha: float
    dt: float
    u0: np.ndarray  # (n+2g, n+2g, 1, 1) float32


def heat_2d_config(seed=0, n=256, alpha=0.01, ghost_cells=1):
    dx = 1.0 / (n - 1)
    dt = 0.2 * dx * dx / alpha
    g = ghost_cells
    coords = np.arange(-g, n + g, dtype=np.float64) * dx
    rng = np.random.default_rng(seed)
    u = gaussian_bumps(rng, coords, coords)
    u0 = u.astype(np.float32).reshape(n + 2 * g, n + 2 * g, 1, 1)
    return Heat2D(n, g, dx, alpha, dt, u0)


def operator_3d_fields(seed=0, n=512, ghost_cells=1):
    """Interior-only draws (n^3 scalar, n^3 x 3 vector) padded by ghost_cells;
    ghosts are left 0 for the caller to fill periodically."""
    rng = np.random.default_rng(seed)
    g = ghost_cells
    P = n + 2 * g
    s = np.zeros((P, P, P, 1), np.float32)
    v = np.zeros((P, P, P, 3), np.float32)
    s[g:P - g, g:P - g, g:P - g, 0] = rng.uniform(-1, 1, size=(n, n, n))
    v[g:P - g, g:P - g, g:P - g, :] = rng.uniform(-1, 1, size=(n, n, n, 3))
    return s, v


def smoothed_field(rng, n, corr):
    """Gaussian random field on an n x n periodic grid: white noise filtered by a
    Gaussian of standard deviation ``corr`` cells (in Fourier space),
    normalised to max |g| = 1."""
    noise = rng.standard_normal((n, n))
    k = np.fft.fftfreq(n)
    kx, ky = np.meshgrid(k, np.fft.rfftfreq(n), indexing="ij")
    filt = np.exp(-2.0 * (np.pi * corr) ** 2 * (kx ** 2 + ky ** 2))
    g = np.fft.irfft2(np.fft.rfft2(noise) * filt, s=(n, n))
    return g / np.max(np.abs(g))


def ricker(t, f0, t0):
    a = (np.pi * f0 * (t - t0)) ** 2
    return (1.0 - 2.0 * a) * np.exp(-a)


@dataclass
class Wave2D:
    n: int
    g: int
    dx: float
    dt: float
    steps: int
    c2: np.ndarray          # (n+2g, n+2g, 1, 1) float32, squared wave speed per node
    src_idx: np.ndarray     # (1,) flat node index (centre)
    wavelet: np.ndarray     # (steps, 1) float32 source samples
    rec_idx: np.ndarray     # (nrec,) flat node indices, row next to the +Y wall
    f0: float


def wave_2d_config(seed=0, n=2048, steps=2000, nrec=64, corr=32.0, ghost_cells=1,
                   perturbation=0.3, rec_distance=None):
    """2D acoustic workload: c = 1 + perturbation * g (g a smoothed Gaussian
    random field, correlation length ``corr`` cells), dt = 0.5 dx / c_max (CFL
    0.5), Ricker source at the centre with f0 = c_min / (20 dx) (20 cells per
    wavelength) and t0 = 1.5 / f0, ``nrec`` receivers evenly spaced on the row
    next to the +Y wall, Dirichlet-0 walls.  ``perturbation=0`` gives the
    unperturbed model (c = 1).

    At CFL 0.5 the front moves at most half a cell per step, so after 2000
    steps on 2048^2 it is at most 1000 cells from the source while the wall
    row is 1022 away: those receivers record exactly 0.  ``rec_distance``
    (cells) puts the receiver row that far from the source along +Y instead,
    where the wave arrives; everything else is unchanged."""
    rng = np.random.default_rng(seed)
    g = ghost_cells
    P = n + 2 * g
    dx = 1.0 / (n - 1)
    field = smoothed_field(rng, n, corr) if perturbation else np.zeros((n, n))
    c = 1.0 + perturbation * field
    cpad = np.pad(c, g, mode="edge")
    c2 = (cpad ** 2).astype(np.float32).reshape(P, P, 1, 1)
    cmax, cmin = float(c.max()), float(c.min())
    dt = 0.5 * dx / cmax
    f0 = cmin / (20.0 * dx)
    t = np.arange(steps) * dt
    wav = ricker(t, f0, 1.5 / f0).astype(np.float32)[:, None]
    src = np.array([(g + n // 2) * P + (g + n // 2)], dtype=np.int64)
    cols = np.linspace(g + 1, g + n - 2, nrec).round().astype(np.int64)
    row = P - 2 - g                                # next to the +Y wall
    if rec_distance is not None:
        row = g + n // 2 + int(rec_distance)
        if not g + n // 2 < row < P - g:
            raise ValueError("receiver row outside the interior")
    rec = cols * P + row
    return Wave2D(n, g, dx, dt, steps, c2, src, wav, rec, f0)


def gaussian_bumps_separable(rng, coords_x, coords_y, n_bumps=8):
    """The same bump field as ``gaussian_bumps`` (same draws, same order),
    each bump evaluated as the outer product exp(-dx^2/2w^2) exp(-dy^2/2w^2):
    O(n) exponentials per bump instead of O(n^2)."""
    centres = rng.uniform(0.2, 0.8, size=(n_bumps, 2))
    widths = rng.uniform(0.02, 0.08, size=n_bumps)
    amps = rng.uniform(0.5, 1.0, size=n_bumps)
    u = np.zeros((len(coords_x), len(coords_y)))
    for (cx, cy), w, a in zip(centres, widths, amps):
        ex = np.exp(-((coords_x - cx) ** 2) / (2.0 * w * w))
        ey = np.exp(-((coords_y - cy) ** 2) / (2.0 * w * w))
        u += a * np.outer(ex, ey)
    return u


@dataclass
class HeatEnsembleConfig:
    g: int
    n: np.ndarray        # interior nodes per side, per instance
    dx: np.ndarray
    alpha: np.ndarray
    dt: np.ndarray
    steps: int

    def u0(self, i, seed):
        """Initial field of instance i: 8 Gaussian bumps from default_rng([seed, i])."""
        n, g = int(self.n[i]), self.g
        coords = np.arange(-g, n + g, dtype=np.float64) * self.dx[i]
        u = gaussian_bumps_separable(np.random.default_rng([seed, i]), coords, coords)
        return u.astype(np.float32).reshape(n + 2 * g, n + 2 * g, 1, 1)


def heat_ensemble_config(seed=3, count=512, n_min=64, n_max=1024, steps=500, ghost_cells=1):
    """Batched heat workload: side lengths rint(exp(U[ln n_min, ln n_max])),
    diffusivities U[0.005, 0.05] (both from default_rng(seed), in that order),
    dx = 1/(n-1), dt = 0.2 dx^2 / alpha, Dirichlet 0 walls."""
    rng = np.random.default_rng(seed)
    n = np.rint(np.exp(rng.uniform(np.log(n_min), np.log(n_max), count))).astype(np.int64)
    alpha = rng.uniform(0.005, 0.05, count)
    dx = 1.0 / (n - 1)
    dt = 0.2 * dx * dx / alpha
    return HeatEnsembleConfig(ghost_cells, n, dx, alpha, dt, steps)


def ns_cavity_u0(seed, n, ghost_cells=1, amplitude=1e-3, corr=8.0):
    """Initial velocity of the projection cavity workload: a smoothed Gaussian
    random streamfunction psi (correlation length ``corr`` cells, seeded), its
    discrete curl u = (d psi / dy, -d psi / dx) by central differences on the
    padded grid (so the discrete divergence of the interior is 0 up to
    rounding), scaled to max |u| = amplitude.  Returns (n+2g, n+2g, 1, 2)."""
    rng = np.random.default_rng(seed)
    g = ghost_cells
    P = n + 2 * g
    psi = smoothed_field(rng, P, corr)
    u = np.zeros((P, P, 2))
    u[1:-1, 1:-1, 0] = (psi[1:-1, 2:] - psi[1:-1, :-2]) / 2.0
    u[1:-1, 1:-1, 1] = -(psi[2:, 1:-1] - psi[:-2, 1:-1]) / 2.0
    u *= amplitude / np.max(np.abs(u))
    return u.astype(np.float32).reshape(P, P, 1, 2)
