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


def _point_set(idx, p2, data, scale):
    e = np.asarray(idx, np.int64)
    rows, cols = e // p2, e % p2
    order = np.argsort(rows, kind="stable")
    return rows, cols, order


def _points(layout, p1, p2):
    """Source and receiver flat indices: a few scattered points ("spread"), or
    rows holding more than 64 points each, duplicates included ("dense_row":
    the fused kernels walk a row's list 64 points at a time)."""
    if layout == "spread":
        src = np.array([5 * p2 + 7, 2 * p2 + 3, 5 * p2 + 7, (p1 - 3) * p2 + p2 - 4])
        rec = np.array([1 * p2 + 1, 6 * p2 + 9, 3 * p2 + 4, (p1 - 2) * p2 + 2])
        return src, rec
    cols = np.arange(70) * (p2 - 1) // 69
    src = np.concatenate([5 * p2 + cols, [5 * p2 + 7, 5 * p2 + 7, 2 * p2 + 3]])
    # columns 255 / 256 straddle the first 256-column strip edge
    rec = np.concatenate([(p1 - 2) * p2 + cols[::-1],
                          (p1 - 2) * p2 + np.array([256, 256, 255]), [6 * p2 + 1]])
    return src, rec


@pytest.mark.parametrize("shape,layout", [((40, 37, 1), "spread"), ((1, 70, 300), "spread"),
                                          ((40, 300, 1), "dense_row")])
def test_fused_wave_steps_match_separate_kernels(dev, shape, layout):
    """pde_wave_step_fused_f32 / pde_wave_step_adjoint_fused_f32 against the
    separate source, step, receiver and counter kernels, over 5 steps and back."""
    print(f"fused wave steps {shape} {layout}", flush=True)
    p1, p2 = [s for s in shape if s > 1]
    rng = np.random.default_rng(9)
    u, v, c2 = (T(x, dev) for x in _fields(shape, 3, 11))
    steps, g = 5, 1
    src, rec = _points(layout, p1, p2)
    wav = T(rng.uniform(-1, 1, (steps, len(src))).astype(np.float32), dev)
    w = O.laplacian_weights(0.1)
    alpha, dt = np.float32(0.7), np.float32(0.05)
    keep = []

    def ps(idx, data, stride, scale):
        rows, cols, order = _point_set(idx, p2, data, scale)
        rs = np.searchsorted(rows[order], np.arange(p1 + 1)).astype(np.int32)
        arrs = [T(rs, dev), T(cols[order].astype(np.int32), dev), T(order.astype(np.int32), dev)]
        keep.extend(arrs)
        r = np.zeros(1, N.POINT_SET_DTYPE)
        r[0] = (arrs[0].data_ptr(), arrs[1].data_ptr(), arrs[2].data_ptr(), data.data_ptr(),
                stride, np.float32(scale), 0)
        keep.append(r)
        return r.ctypes.data

    st = N.stream_of(u)
    # fused path
    fu, fv = u.clone(), v.clone()
    fo = torch.empty_like(u)
    frec = torch.zeros(steps, len(rec), device=dev)
    ctr = torch.zeros(2, dtype=torch.int64, device=dev)
    ps_src, ps_rec = ps(src, wav, len(src), dt), ps(rec, frec, len(rec), 1.0)
    slot = 0
    for n in range(steps):
        N.call("pde_wave_step_fused_f32", fu.data_ptr(), fv.data_ptr(), c2.data_ptr(),
               fo.data_ptr(), fv.data_ptr(), *shape, g, *w, alpha, dt,
               ctr.data_ptr() + 8 * slot, ctr.data_ptr() + 8 * (1 - slot), ps_src, ps_rec, st)
        fu, fo = fo, fu
        slot = 1 - slot
    # separate kernels (distinct source indices there: the duplicate is summed
    # by two single-source calls in list order)
    su, sv = u.clone(), v.clone()
    so = torch.empty_like(u)
    srec = torch.zeros(steps, len(rec), device=dev)
    c1 = torch.zeros(1, dtype=torch.int64, device=dev)
    sidx = T(src, dev)
    ridx = T(rec, dev)
    for n in range(steps):
        for j in range(len(src)):
            N.call("pde_scatter_add_step_f32", sv.data_ptr(), sidx[j:].data_ptr(),
                   wav[:, j:].data_ptr(), 1, len(src), dt, c1.data_ptr(), st)
        N.call("pde_wave_step_ctr_f32", su.data_ptr(), sv.data_ptr(), c2.data_ptr(),
               so.data_ptr(), sv.data_ptr(), *shape, g, *w, alpha, dt, c1.data_ptr(), st)
        N.call("pde_gather_step_f32", so.data_ptr(), ridx.data_ptr(), srec.data_ptr(), len(rec),
               len(rec), c1.data_ptr(), st)
        N.call("pde_counter_add", c1.data_ptr(), 1, st)
        su, so = so, su
    assert same_bits(fu.cpu().numpy(), su.cpu().numpy())
    assert same_bits(fv.cpu().numpy(), sv.cpu().numpy())
    assert same_bits(frec.cpu().numpy(), srec.cpu().numpy())
    assert int(ctr[slot]) == steps == int(c1[0])

    # transposes, one step from n = steps - 1
    gun, gvn, u2 = (T(x, dev) for x in _fields(shape, 3, 12))
    grec = T(rng.uniform(-1, 1, (steps, len(rec))).astype(np.float32), dev)
    outs = [torch.empty_like(u) for _ in range(4)]
    gc_f, gc_s = torch.zeros_like(u), torch.zeros_like(u)
    ctr.fill_(0)
    ctr[1] = steps - 1
    ps_grec = ps(rec, grec, len(rec), 1.0)
    N.call("pde_wave_step_adjoint_fused_f32", gun.data_ptr(), gvn.data_ptr(), c2.data_ptr(),
           u2.data_ptr(), outs[0].data_ptr(), outs[1].data_ptr(), gc_f.data_ptr(), *shape, g, *w,
           alpha, dt, ctr.data_ptr() + 8, ctr.data_ptr(), ps_grec, st)
    c1.fill_(steps - 1)
    gun2 = gun.clone()
    for j in range(len(rec)):     # one point per call: duplicates summed in list order
        N.call("pde_scatter_add_step_f32", gun2.data_ptr(), ridx[j:].data_ptr(),
               grec[:, j:].data_ptr(), 1, len(rec), np.float32(1.0), c1.data_ptr(), st)
    N.call("pde_wave_step_adjoint_ctr_f32", gun2.data_ptr(), gvn.data_ptr(), c2.data_ptr(),
           u2.data_ptr(), outs[2].data_ptr(), outs[3].data_ptr(), gc_s.data_ptr(), *shape, g, *w,
           alpha, dt, c1.data_ptr(), st)
    assert same_bits(outs[0].cpu().numpy(), outs[2].cpu().numpy())
    assert same_bits(outs[1].cpu().numpy(), outs[3].cpu().numpy())
    assert same_bits(gc_f.cpu().numpy(), gc_s.cpu().numpy())
    assert int(ctr[0]) == steps - 2
