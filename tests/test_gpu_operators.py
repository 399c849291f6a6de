"""GPU parity of the operator kernels against the C FP32 restatement: bit for
bit on every non-ghost node, output ghosts left untouched."""
import numpy as np
import pytest
import torch

from conftest import same_bits
from helpers import ghost_free_mask
from oracle import np_oracle as O
from pde_module_amd import Divergence, Grad, Laplacian, PeriodicBoundary, PeriodicLaplacian
import pde_module_amd._native as N
from pde_module_amd.FDM import functional as F

pytestmark = pytest.mark.gpu

SHAPES = [(10, 12, 1), (6, 7, 8), (9, 1, 1), (1, 6, 5), (33, 17, 70), (130, 9, 1), (5, 70, 66)]


def to_dev(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("C", [1, 3])
@pytest.mark.parametrize("g", [1, 2])
def test_laplacian_matches_oracle(dev, corc, shape, C, g):
    print(f"laplacian shape={shape} C={C} g={g}", flush=True)
    rng = np.random.default_rng(7)
    u = rng.uniform(-1, 1, shape + (C,)).astype(np.float32)
    dx, alpha = 0.037, 0.6
    lap = Laplacian(C, dx, g)
    got = lap(to_dev(u, dev), alpha).cpu().numpy()
    want = corc.laplacian(u, lap.stencil, np.float32(alpha), g)
    assert same_bits(got, want, ghost_free_mask(shape, g, (C,)))


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("C", [1, 2])
def test_grad_and_divergence_match_oracle(dev, corc, shape, C):
    print(f"grad/div shape={shape} C={C}", flush=True)
    rng = np.random.default_rng(8)
    g, dx = 1, 0.05
    u = rng.uniform(-1, 1, shape + (C,)).astype(np.float32)
    D = len(O.eligible_axes(shape))
    grad = Grad(C, shape, dx, g)
    got = grad(to_dev(u, dev), 1.7).cpu().numpy()
    want = corc.grad(u, grad.stencil, np.float32(1.7), g)
    assert got.shape == want.shape
    mask = ghost_free_mask(shape, g, got.shape[3:])
    assert same_bits(got, want, mask)
    # vector divergence (vec D -> vec 1) and matrix divergence (C x D -> vec C)
    v = rng.uniform(-1, 1, shape + (D,)).astype(np.float32)
    div = Divergence(D, shape, dx, g)
    got = div(to_dev(v, dev), -0.3).cpu().numpy()
    want = corc.divergence(v, div.stencil, np.float32(-0.3), g)
    assert same_bits(got, want, ghost_free_mask(shape, g, (1,)))
    m = rng.uniform(-1, 1, shape + (C, D)).astype(np.float32)
    divm = Divergence((C, D), shape, dx, g)
    got = divm(to_dev(m, dev), 2.5).cpu().numpy()
    want = corc.divergence(m, divm.stencil, np.float32(2.5), g, matrix_in=True)
    assert same_bits(got, want, ghost_free_mask(shape, g, (C,)))


def test_grad_force_matrix_layout(dev, corc):
    rng = np.random.default_rng(9)
    shape = (12, 11, 1)
    u = rng.uniform(-1, 1, shape + (1,)).astype(np.float32)
    gm = Grad(1, shape, 0.1, 1, force_matrix=True)
    got = gm(to_dev(u, dev)).cpu().numpy()
    assert got.shape == shape + (1, 2)
    want = corc.grad(u, gm.stencil, np.float32(1.0), 1, force_matrix=True)
    assert same_bits(got, want, ghost_free_mask(shape, 1, (1, 2)))


def test_output_ghosts_untouched(dev):
    shape, g = (20, 18, 9), 2
    u = torch.rand(shape + (1,), device=dev)
    out = torch.full(shape + (1,), float("nan"), device=dev)
    k = Laplacian(1, 0.1, g)
    k.setup(u)
    F.laplacian(k.kernel, u, 1.0, out, g)
    o = out.cpu().numpy()
    m = ghost_free_mask(shape, g, (1,))
    assert np.all(np.isnan(o[~m])) and not np.any(np.isnan(o[m]))


@pytest.mark.parametrize("shape", [(131, 67, 200), (20, 19, 1030), (5, 3, 4), (34, 9, 517),
                                   (70, 33, 515), (3, 3, 3), (9, 17, 1), (10, 6, 4), (10, 6, 6),
                                   (7, 13, 1030), (12, 6, 514), (9, 7, 7), (6, 11, 516)])
def test_tiled_kernel_equals_generic_kernel(dev, shape):
    """The 3D tiled Laplacian (partial lanes, a last row tile shifted onto the
    one before, several z tiles, z tiles of 1..4 points, short x chunks)
    against the one-thread-per-point kernel."""
    print(f"tiled vs generic {shape}", flush=True)
    u = torch.rand(shape + (1,), device=dev) * 2 - 1
    a = torch.zeros_like(u)
    b = torch.zeros_like(u)
    s = N.stream_of(u)
    w = (np.float32(4.0), np.float32(-8.0), np.float32(4.0))
    N.call("pde_laplacian_f32", u.data_ptr(), a.data_ptr(), *shape, 1, 1, *w, np.float32(0.3), s)
    N.call("pde_laplacian_generic_f32", u.data_ptr(), b.data_ptr(), *shape, 1, 1, *w,
           np.float32(0.3), s)
    assert torch.equal(a.view(torch.int32), b.view(torch.int32))


def test_fused_pass_equals_separate_modules(dev):
    shape, g, dx = (40, 37, 70), 1, 1.0 / 64
    s = torch.rand(shape + (1,), device=dev) * 2 - 1
    v = torch.rand(shape + (3,), device=dev) * 2 - 1
    lap = torch.zeros_like(s)
    grad = torch.zeros_like(v)
    div = torch.zeros_like(s)
    F.fused_lap_grad_div(s, v, lap, grad, div, dx, g, 0.5, 1.5, -2.0)
    L = Laplacian(1, dx, g)(s, 0.5)
    G = Grad(1, shape, dx, g)(s, 1.5)
    D = Divergence(3, shape, dx, g)(v, -2.0)
    m = torch.from_numpy(ghost_free_mask(shape, g, ())).to(dev)
    for got, want in ((lap, L), (grad, G), (div, D)):
        assert torch.equal(got[m].view(torch.int32), want[m].view(torch.int32))


@pytest.mark.parametrize("shape,g", [((10, 11, 12), 1), ((9, 10, 1), 2), ((8, 1, 1), 1),
                                     ((12, 9, 13), 3), ((7, 300, 5), 1)])
def test_periodic_fill_matches_oracle(dev, corc, shape, g):
    rng = np.random.default_rng(4)
    u = rng.uniform(-1, 1, shape + (2,)).astype(np.float32)
    got = PeriodicBoundary(g)(to_dev(u, dev)).cpu().numpy()
    want = corc.periodic_fill(u.copy(), g)
    assert same_bits(got, want)


BAND_SHAPES = [((12, 9, 517), 1), ((7, 13, 1030), 1), ((10, 6, 6), 1), ((10, 6, 4), 1),
               ((9, 12, 11), 2), ((40, 37, 70), 1), ((6, 4, 9), 1)]


@pytest.mark.parametrize("shape,g", BAND_SHAPES)
def test_band_grad_div_fused_match_oracle(dev, corc, shape, g):
    """3D scalar Grad, vec3 Divergence and the fused lap+grad+div pass through
    the band kernels (partial lanes, several z tiles, z tiles of 1..4 points,
    a shifted last row tile, g = 2) against the C restatement."""
    print(f"band grad/div/fused shape={shape} g={g}", flush=True)
    rng = np.random.default_rng(12)
    dx = 0.029
    s = rng.uniform(-1, 1, shape + (1,)).astype(np.float32)
    v = rng.uniform(-1, 1, shape + (3,)).astype(np.float32)
    grad = Grad(1, shape, dx, g)
    got_g = grad(to_dev(s, dev), 1.3).cpu().numpy()
    want_g = corc.grad(s, grad.stencil, np.float32(1.3), g)
    assert same_bits(got_g, want_g, ghost_free_mask(shape, g, (3,)))
    div = Divergence(3, shape, dx, g)
    got_d = div(to_dev(v, dev), -0.8).cpu().numpy()
    want_d = corc.divergence(v, div.stencil, np.float32(-0.8), g)
    assert same_bits(got_d, want_d, ghost_free_mask(shape, g, (1,)))
    lap = Laplacian(1, dx, g)
    want_l = corc.laplacian(s, lap.stencil, np.float32(0.5), g)
    sd, vd = to_dev(s, dev), to_dev(v, dev)
    lo = torch.full_like(sd, float("nan"))
    go = torch.full_like(vd, float("nan"))
    do = torch.full_like(sd, float("nan"))
    F.fused_lap_grad_div(sd, vd, lo, go, do, dx, g, 0.5, 1.3, -0.8)
    assert same_bits(lo.cpu().numpy(), want_l, ghost_free_mask(shape, g, (1,)))
    assert same_bits(go.cpu().numpy(), want_g, ghost_free_mask(shape, g, (3,)))
    assert same_bits(do.cpu().numpy(), want_d, ghost_free_mask(shape, g, (1,)))
    m = ghost_free_mask(shape, g, (1,))
    assert np.all(np.isnan(lo.cpu().numpy()[~m]))


PERIODIC_SHAPES = [((10, 11, 12), 1, 1), ((34, 9, 517), 1, 1), ((7, 13, 1030), 1, 1),
                   ((12, 10, 9), 2, 1), ((10, 6, 4), 1, 1), ((5, 3, 4), 1, 1), ((10, 12, 1), 1, 1),
                   ((9, 10, 1), 2, 3), ((1, 6, 5), 1, 2), ((8, 1, 1), 1, 1), ((11, 9, 8), 1, 3)]


@pytest.mark.parametrize("shape,g,C", PERIODIC_SHAPES)
def test_periodic_laplacian_matches_fill_then_laplacian(dev, corc, shape, g, C):
    """PeriodicLaplacian == periodic ghost fill + Laplacian (the C oracle), bit
    for bit on the interior; the input (ghosts included) and the output's
    ghosts are left untouched."""
    print(f"periodic laplacian shape={shape} g={g} C={C}", flush=True)
    rng = np.random.default_rng(11)
    u = rng.uniform(-1, 1, shape + (C,)).astype(np.float32)
    dx, alpha = 0.041, -0.7
    lap = PeriodicLaplacian(C, dx, g)
    ud = to_dev(u, dev)
    lap.setup(ud)
    lap.output_array.fill_(float("nan"))
    got = lap(ud, alpha).cpu().numpy()
    filled = corc.periodic_fill(u.copy(), g)
    want = corc.laplacian(filled, lap.stencil, np.float32(alpha), g)
    mask = ghost_free_mask(shape, g, (C,))
    assert same_bits(got, want, mask)
    assert np.all(np.isnan(got[~mask]))
    assert same_bits(ud.cpu().numpy(), u)


# Interior extents whose rows split into whole 256-point wave segments run the
# LDS plane-ring kernel: one and two waves per row segment, several z tiles,
# partial and single-plane x chunks, a shifted last row tile, ghost widths 1-3.
RING_SHAPES = [((13, 5, 256), 1), ((25, 9, 512), 1), ((2, 4, 768), 2), ((14, 7, 1024), 3),
               ((1, 4, 256), 1), ((24, 6, 512), 2), ((3, 11, 512), 1)]


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("interior,g", RING_SHAPES)
def test_ring_laplacian_matches_oracle(dev, corc, interior, g, periodic):
    """The 3D scalar Laplacian on grids the plane-ring kernel takes, plain and
    periodic, against the C restatement bit for bit; ghosts of the output and
    the whole input left untouched."""
    shape = tuple(e + 2 * g for e in interior)
    print(f"ring laplacian shape={shape} g={g} periodic={periodic}", flush=True)
    rng = np.random.default_rng(21)
    u = rng.uniform(-1, 1, shape + (1,)).astype(np.float32)
    dx, alpha = 0.043, 0.9
    lap = (PeriodicLaplacian if periodic else Laplacian)(1, dx, g)
    ud = to_dev(u, dev)
    lap.setup(ud)
    lap.output_array.fill_(float("nan"))
    got = lap(ud, alpha).cpu().numpy()
    src = corc.periodic_fill(u.copy(), g) if periodic else u
    want = corc.laplacian(src, lap.stencil, np.float32(alpha), g)
    mask = ghost_free_mask(shape, g, (1,))
    assert same_bits(got, want, mask)
    assert np.all(np.isnan(got[~mask]))
    assert same_bits(ud.cpu().numpy(), u)


@pytest.mark.parametrize("interior,g", RING_SHAPES)
def test_ring_divergence_matches_oracle(dev, corc, interior, g):
    """vec3 Divergence on grids the Divergence ring takes (one wave per
    256-point row segment, 16-plane chunks marched up and down, a shifted
    last row tile) against the C restatement bit for bit; output ghosts
    untouched."""
    shape = tuple(e + 2 * g for e in interior)
    print(f"ring divergence shape={shape} g={g}", flush=True)
    rng = np.random.default_rng(22)
    v = rng.uniform(-1, 1, shape + (3,)).astype(np.float32)
    div = Divergence(3, shape, 0.047, g)
    vd = to_dev(v, dev)
    div.setup(vd)
    div.output_array.fill_(float("nan"))
    got = div(vd, -1.1).cpu().numpy()
    want = corc.divergence(v, div.stencil, np.float32(-1.1), g)
    mask = ghost_free_mask(shape, g, (1,))
    assert same_bits(got, want, mask)
    assert np.all(np.isnan(got[~mask]))


@pytest.mark.parametrize("interior,g", [((25, 9, 512), 1), ((3, 11, 512), 1), ((14, 8, 256), 2),
                                        ((17, 12, 768), 1)])
def test_split_fused_pass_matches_oracle(dev, corc, interior, g):
    """The fused Laplacian + Grad + Divergence step on grids where it runs as
    a scalar band launch and the Divergence ring, against the C restatement
    bit for bit."""
    shape = tuple(e + 2 * g for e in interior)
    print(f"split fused shape={shape} g={g}", flush=True)
    rng = np.random.default_rng(23)
    dx = 0.031
    s = rng.uniform(-1, 1, shape + (1,)).astype(np.float32)
    v = rng.uniform(-1, 1, shape + (3,)).astype(np.float32)
    sd, vd = to_dev(s, dev), to_dev(v, dev)
    lo = torch.full_like(sd, float("nan"))
    go = torch.full_like(vd, float("nan"))
    do = torch.full_like(sd, float("nan"))
    F.fused_lap_grad_div(sd, vd, lo, go, do, dx, g, 0.5, 1.3, -0.8)
    lap, grad, div = Laplacian(1, dx, g), Grad(1, shape, dx, g), Divergence(3, shape, dx, g)
    assert same_bits(lo.cpu().numpy(), corc.laplacian(s, lap.stencil, np.float32(0.5), g),
                     ghost_free_mask(shape, g, (1,)))
    assert same_bits(go.cpu().numpy(), corc.grad(s, grad.stencil, np.float32(1.3), g),
                     ghost_free_mask(shape, g, (3,)))
    assert same_bits(do.cpu().numpy(), corc.divergence(v, div.stencil, np.float32(-0.8), g),
                     ghost_free_mask(shape, g, (1,)))


def test_ring_kernels_every_chunk_length(dev, corc):
    """Every x-chunk length of the plane rings (1 .. 17 interior planes: single
    and partial chunks, chunks shorter than the planes in flight, up and down
    marches) for the plain and periodic Laplacian and the vec3 Divergence,
    against the C restatement bit for bit."""
    rng = np.random.default_rng(24)
    dx = 0.039
    for e0 in range(1, 18):
        shape = (e0 + 2, 6, 258)
        print(f"ring chunk sweep shape={shape}", flush=True)
        u = rng.uniform(-1, 1, shape + (1,)).astype(np.float32)
        ud = to_dev(u, dev)
        for periodic in (False, True):
            lap = (PeriodicLaplacian if periodic else Laplacian)(1, dx, 1)
            got = lap(ud, 0.8).cpu().numpy()
            src = corc.periodic_fill(u.copy(), 1) if periodic else u
            want = corc.laplacian(src, lap.stencil, np.float32(0.8), 1)
            assert same_bits(got, want, ghost_free_mask(shape, 1, (1,))), (e0, periodic)
        v = rng.uniform(-1, 1, shape + (3,)).astype(np.float32)
        div = Divergence(3, shape, dx, 1)
        got = div(to_dev(v, dev), 1.2).cpu().numpy()
        want = corc.divergence(v, div.stencil, np.float32(1.2), 1)
        assert same_bits(got, want, ghost_free_mask(shape, 1, (1,))), e0


def test_full_size_periodic_laplacian_matches_oracle(dev, corc):
    """The 512^3 periodic benchmark grid (514^3 padded) against the OpenMP C
    restatement, bit for bit on all 134M interior points."""
    n, g = 512, 1
    P = n + 2 * g
    print("full-size 514^3 laplacian", flush=True)
    gen = torch.Generator(device=dev).manual_seed(1)
    u = torch.rand((P, P, P, 1), device=dev, generator=gen) * 2 - 1
    PeriodicBoundary(g)(u)
    lap = Laplacian(1, 1.0 / n, g)
    got = lap(u, 1.0).cpu().numpy()
    uh = u.cpu().numpy()
    want = corc.laplacian(uh, lap.stencil, np.float32(1.0), g, nthreads=min(16, corc.max_threads()))
    inner = (slice(1, -1),) * 3
    assert np.array_equal(got[inner].view(np.uint32), want[inner].view(np.uint32))
    # the wrapped-index kernel on the same grid, ghosts left stale
    u[0].zero_()
    u[:, -1].zero_()
    u[:, :, 0].zero_()
    per = PeriodicLaplacian(1, 1.0 / n, g)(u, 1.0).cpu().numpy()
    assert np.array_equal(per[inner].view(np.uint32), want[inner].view(np.uint32))


@pytest.mark.parametrize("periodic", [False, True])
def test_long_chunk_ring_laplacian_matches_oracle(dev, corc, periodic):
    """A grid large enough for the Laplacian ring's 32-plane chunks (233 x 1024
    x 512 interior: 8 chunks, the last of 9 planes marching down) against the
    OpenMP C restatement, bit for bit."""
    g = 1
    shape = (233 + 2, 1024 + 2, 512 + 2)
    print(f"long-chunk ring laplacian shape={shape} periodic={periodic}", flush=True)
    gen = torch.Generator(device=dev).manual_seed(5)
    u = torch.rand(shape + (1,), device=dev, generator=gen) * 2 - 1
    lap = (PeriodicLaplacian if periodic else Laplacian)(1, 0.01, g)
    got = lap(u, 0.9).cpu().numpy()
    uh = u.cpu().numpy()
    nt = min(16, corc.max_threads())
    src = corc.periodic_fill(uh.copy(), g) if periodic else uh
    want = corc.laplacian(src, lap.stencil, np.float32(0.9), g, nthreads=nt)
    inner = (slice(1, -1),) * 3
    assert np.array_equal(got[inner].view(np.uint32), want[inner].view(np.uint32))


def test_full_size_grad_div_fused_match_oracle(dev, corc):
    """The 512^3 operator workload (514^3 padded, scalar s and vec3 v ~ U[-1,1]):
    Grad (8-row band tiles), Divergence and the fused lap+grad+div pass against
    the OpenMP C restatement, bit for bit on all 134M interior points."""
    n, g = 512, 1
    P = n + 2 * g
    dx = 1.0 / n
    print("full-size 514^3 grad / div / fused", flush=True)
    gen = torch.Generator(device=dev).manual_seed(2)
    s = torch.rand((P, P, P, 1), device=dev, generator=gen) * 2 - 1
    v = torch.rand((P, P, P, 3), device=dev, generator=gen) * 2 - 1
    nt = min(16, corc.max_threads())
    inner = (slice(1, -1),) * 3
    sh, vh = s.cpu().numpy(), v.cpu().numpy()
    eq = lambda a, b: np.array_equal(a[inner].view(np.uint32), b[inner].view(np.uint32))
    grad = Grad(1, (P, P, P), dx, g)
    want_g = corc.grad(sh, grad.stencil, np.float32(1.0), g, nthreads=nt)
    assert eq(grad(s, 1.0).cpu().numpy(), want_g)
    div = Divergence(3, (P, P, P), dx, g)
    want_d = corc.divergence(vh, div.stencil, np.float32(1.0), g, nthreads=nt)
    assert eq(div(v, 1.0).cpu().numpy(), want_d)
    lap = Laplacian(1, dx, g)
    want_l = corc.laplacian(sh, lap.stencil, np.float32(1.0), g, nthreads=nt)
    lo, do = torch.empty_like(s), torch.empty_like(s)
    go = torch.empty_like(v)
    F.fused_lap_grad_div(s, v, lo, go, do, dx, g, 1.0, 1.0, 1.0)
    assert eq(lo.cpu().numpy(), want_l)
    assert eq(go.cpu().numpy(), want_g)
    assert eq(do.cpu().numpy(), want_d)


def test_long_chunk_grad_div_match_oracle(dev, corc):
    """Grad and Divergence on a grid large enough for Divergence's 16-plane
    chunks (interior 118 x 510 x 520: two z tiles, the second of 8 points,
    64 row tiles with a shifted last one, 8 chunks of which the last has 6
    planes; 8 x 16 x 8 = 1024 blocks) and Grad's 2-plane chunks, against the
    OpenMP C restatement, bit for bit."""
    shape, g, dx = (120, 512, 522), 1, 0.013
    print(f"long-chunk grad/div {shape}", flush=True)
    gen = torch.Generator(device=dev).manual_seed(5)
    s = torch.rand(shape + (1,), device=dev, generator=gen) * 2 - 1
    v = torch.rand(shape + (3,), device=dev, generator=gen) * 2 - 1
    nt = min(16, corc.max_threads())
    grad = Grad(1, shape, dx, g)
    want_g = corc.grad(s.cpu().numpy(), grad.stencil, np.float32(0.7), g, nthreads=nt)
    assert same_bits(grad(s, 0.7).cpu().numpy(), want_g, ghost_free_mask(shape, g, (3,)))
    div = Divergence(3, shape, dx, g)
    want_d = corc.divergence(v.cpu().numpy(), div.stencil, np.float32(-1.1), g, nthreads=nt)
    assert same_bits(div(v, -1.1).cpu().numpy(), want_d, ghost_free_mask(shape, g, (1,)))
