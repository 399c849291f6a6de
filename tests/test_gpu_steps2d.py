"""GPU parity of the 2D scalar row-tile step kernels (fused heat and wave steps
and their transposes) against the NumPy FP32 oracle, bit for bit, on shapes
that exercise every edge of the tiling: rows not a multiple of the 16-row
block, columns not a multiple of 4 (partial 16-byte chunks) or of the
256-column wave strip, a strip with a single chunk, tiny grids, ghost widths 1
and 2, and each placement of the size-1 axis."""
import numpy as np
import pytest
import torch

from conftest import same_bits
from oracle import np_oracle as O
import pde_module_amd._native as N

pytestmark = pytest.mark.gpu

SHAPES = [
    ((3, 3, 1), 1),
    ((5, 7, 1), 1),
    ((17, 257, 1), 1),
    ((33, 260, 1), 2),
    ((18, 1, 513), 1),
    ((1, 40, 1030), 2),
    ((7, 2051, 1), 1),
    ((130, 258, 1), 1),
]


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _fields(shape, n, seed):
    rng = np.random.default_rng(seed)
    return [rng.uniform(-1, 1, shape + (1,)).astype(np.float32) for _ in range(n)]


@pytest.mark.parametrize("shape,g", SHAPES)
def test_heat_step_tile(dev, shape, g):
    print(f"heat 2d tile {shape} g={g}", flush=True)
    u2, = _fields(shape, 1, 1)
    w = O.laplacian_weights(0.1)
    alpha, dt = np.float32(0.3), np.float32(0.01)
    want = O.heat_step(u2, w, alpha, g, dt, lambda x: x)
    M = O.interior_mask(shape, g)[..., None]
    want = np.where(M, want, u2)
    ud = T(u2, dev)
    out = torch.full_like(ud, float("nan"))
    N.call("pde_heat_step_f32", ud.data_ptr(), out.data_ptr(), *shape, 1, g, *w, alpha, dt,
           N.stream_of(ud))
    assert same_bits(out.cpu().numpy(), want)

    gout = torch.full_like(ud, float("nan"))
    N.call("pde_heat_step_adjoint_f32", ud.data_ptr(), gout.data_ptr(), *shape, 1, g, *w, alpha,
           dt, N.stream_of(ud))
    assert same_bits(gout.cpu().numpy(), O.heat_step_adjoint(u2, w, alpha, g, dt))


@pytest.mark.parametrize("shape,g", SHAPES)
@pytest.mark.parametrize("first", [True, False])
@pytest.mark.parametrize("with_c2", [True, False])
def test_wave_step_tile(dev, shape, g, first, with_c2):
    print(f"wave 2d tile {shape} g={g} first={first} c2={with_c2}", flush=True)
    u2, v = _fields(shape, 2, 2)
    rng = np.random.default_rng(5)
    c2 = (1 + 0.3 * rng.uniform(-1, 1, shape + (1,))).astype(np.float32) if with_c2 else None
    w = O.laplacian_weights(0.1)
    alpha, dt = np.float32(0.7), np.float32(0.05)
    un, vn = O.wave_step_core(u2, v, c2, w, alpha, g, dt, first)
    ud, vd = T(u2, dev), T(v, dev)
    c2d = T(c2, dev) if with_c2 else None
    oun = torch.full_like(ud, float("nan"))
    # v_next written over v (the in-place form the loops use)
    N.call("pde_wave_step_f32", ud.data_ptr(), vd.data_ptr(),
           c2d.data_ptr() if c2d is not None else None, oun.data_ptr(), vd.data_ptr(), *shape, g,
           *w, alpha, dt, 1 if first else 0, N.stream_of(ud))
    assert same_bits(oun.cpu().numpy(), un)
    assert same_bits(vd.cpu().numpy(), vn)


@pytest.mark.parametrize("shape,g", SHAPES)
@pytest.mark.parametrize("first", [True, False])
@pytest.mark.parametrize("with_c2,with_gc2", [(True, True), (False, True), (True, False)])
def test_wave_adjoint_tile(dev, shape, g, first, with_c2, with_gc2):
    print(f"wave adjoint 2d tile {shape} g={g} first={first} c2={with_c2} gc2={with_gc2}",
          flush=True)
    gun, gvn, u2, gc2_0 = _fields(shape, 4, 3)
    rng = np.random.default_rng(6)
    c2 = (1 + 0.3 * rng.uniform(-1, 1, shape + (1,))).astype(np.float32) if with_c2 else None
    w = O.laplacian_weights(0.1)
    alpha, dt = np.float32(0.7), np.float32(0.05)
    gu2, gv, dc2 = O.wave_step_adjoint(gun, gvn, c2, u2, w, alpha, g, dt, first)
    d = {k: T(x, dev) for k, x in dict(gun=gun, gvn=gvn, u2=u2, gc2=gc2_0).items()}
    c2d = T(c2, dev) if with_c2 else None
    ogu2 = torch.full_like(d["gun"], float("nan"))
    ogv = torch.full_like(d["gun"], float("nan"))
    N.call("pde_wave_step_adjoint_f32", d["gun"].data_ptr(), d["gvn"].data_ptr(),
           c2d.data_ptr() if c2d is not None else None, d["u2"].data_ptr(), ogu2.data_ptr(),
           ogv.data_ptr(), d["gc2"].data_ptr() if with_gc2 else None, *shape, g, *w, alpha, dt,
           1 if first else 0, N.stream_of(ogu2))
    assert same_bits(ogu2.cpu().numpy(), gu2)
    assert same_bits(ogv.cpu().numpy(), gv)
    want_gc2 = gc2_0 + dc2 if with_gc2 else gc2_0
    assert same_bits(d["gc2"].cpu().numpy(), want_gc2)
