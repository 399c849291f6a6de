"""Host logic of the batched heat ensemble: row-segment choice, tile lists
(every node of every grid covered exactly once, nothing out of bounds) and the
ensemble workload generator."""
import numpy as np
import pytest

from pde_module_amd.batch import build_tiles, segment_width
from pde_module_amd.workloads import gaussian_bumps, gaussian_bumps_separable, heat_ensemble_config


@pytest.mark.parametrize("nq,cw", [(64, 64), (128, 64), (32, 32), (96, 32), (17, 32),
                                   (1, 16), (16, 16), (33, 16), (257, 16)])
def test_segment_width(nq, cw):
    assert segment_width(nq) == cw


@pytest.mark.parametrize("rows,waves", [(1, 1), (2, 4), (4, 4), (8, 2)])
def test_tiles_cover_every_node_once(rows, waves):
    p1 = np.array([3, 66, 130, 1026, 17, 258])
    p2 = np.array([3, 66, 1026, 130, 515, 258])
    t = build_tiles(p1, p2, rows, waves)
    assert t.dtype == np.int32 and t.shape[1] == 4
    seen = [np.zeros((a, b), np.int32) for a, b in zip(p1, p2)]
    for gi, x0, q0, cw in t:
        if gi < 0:
            continue
        assert cw in (16, 32, 64)
        nrow = waves * (64 // cw) * rows
        P1, P2 = p1[gi], p2[gi]
        nq = (P2 + 3) // 4
        assert 0 <= x0 < P1 and 0 <= q0 < nq
        r1 = min(x0 + nrow, P1)
        c0, c1 = q0 * 4, min((q0 + cw) * 4, P2)
        seen[gi][x0:r1, c0:c1] += 1
    for s in seen:
        assert np.all(s == 1)
    # each column strip's tiles sit at block indices congruent mod 8 (one XCD)
    blocks = {}
    for b, (gi, x0, q0, cw) in enumerate(t):
        if gi >= 0:
            blocks.setdefault((gi, q0), []).append((b, x0))
    assert all(len({b % 8 for b, _ in v}) == 1 for v in blocks.values())
    assert (t[:, 0] < 0).sum() < len(t) // 2
    for v in blocks.values():
        assert [x for _, x in v] == sorted(x for _, x in v)


def test_ensemble_workload():
    c = heat_ensemble_config(seed=3)
    assert len(c.n) == 512 and c.n.min() >= 64 and c.n.max() <= 1024
    assert abs((c.n.astype(np.int64) ** 2).sum() / 1e6 - 89.28) < 0.01
    assert np.all((c.alpha >= 0.005) & (c.alpha <= 0.05))
    np.testing.assert_allclose(c.dt, 0.2 * c.dx ** 2 / c.alpha)
    u = c.u0(5, 3)
    assert u.shape == (c.n[5] + 2, c.n[5] + 2, 1, 1) and u.dtype == np.float32


def test_separable_bumps_equal_direct():
    for seed in range(3):
        xs = np.arange(-1, 98) / 96.0
        a = gaussian_bumps(np.random.default_rng(seed), xs, xs)
        b = gaussian_bumps_separable(np.random.default_rng(seed), xs, xs)
        np.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-300)
