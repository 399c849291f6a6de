"""GPU parity of the transpose kernels against the NumPy FP32 restatement of
the transposes (bit for bit), and GPU dot-product checks of the boundary
transpose."""
import numpy as np
import pytest
import torch

from conftest import same_bits
from oracle import np_oracle as O
import pde_module_amd._native as N
from pde_module_amd.FDM import GridBoundary
from pde_module_amd.deprecated import GridBoundary as LegacyGridBoundary

pytestmark = pytest.mark.gpu

SHAPES = [(10, 12, 1), (6, 7, 8), (9, 1, 1), (33, 17, 40)]


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def call(name, *args):
    N.call(name, *args)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("C", [1, 3])
@pytest.mark.parametrize("g", [1, 2])
def test_laplacian_adjoint_matches_oracle(dev, shape, C, g):
    print(f"lap adjoint {shape} C={C} g={g}", flush=True)
    rng = np.random.default_rng(1)
    y = rng.uniform(-1, 1, shape + (C,)).astype(np.float32)
    w = O.laplacian_weights(0.07)
    want = O.laplacian_adjoint(y, w, 0.9, g)
    yd = T(y, dev)
    out = torch.full_like(yd, float("nan"))
    call("pde_laplacian_adjoint_f32", yd.data_ptr(), out.data_ptr(), *shape, C, g, *w,
         np.float32(0.9), N.stream_of(yd))
    assert same_bits(out.cpu().numpy(), want)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("Nc", [1, 2])
def test_grad_and_div_adjoints_match_oracle(dev, shape, Nc):
    print(f"grad/div adjoint {shape} N={Nc}", flush=True)
    rng = np.random.default_rng(2)
    g = 1
    D = len(O.eligible_axes(shape))
    w = O.central_weights(0.05)
    y = rng.uniform(-1, 1, shape + (Nc, D)).astype(np.float32)
    want = O.grad_adjoint(y, w, 1.3, g)
    yd = T(y, dev)
    out = torch.full(shape + (Nc,), float("nan"), device=dev)
    call("pde_grad_adjoint_f32", yd.data_ptr(), out.data_ptr(), *shape, Nc, g, *w,
         np.float32(1.3), N.stream_of(yd))
    assert same_bits(out.cpu().numpy(), want)
    z = rng.uniform(-1, 1, shape + (Nc,)).astype(np.float32)
    want = O.divergence_adjoint(z, w, -0.6, g, D)
    zd = T(z, dev)
    out = torch.full(shape + (Nc, D), float("nan"), device=dev)
    call("pde_divergence_adjoint_f32", zd.data_ptr(), out.data_ptr(), *shape, Nc, g, *w,
         np.float32(-0.6), N.stream_of(zd))
    assert same_bits(out.cpu().numpy(), want)


@pytest.mark.parametrize("shape", [(20, 18, 1), (9, 10, 11)])
def test_heat_step_adjoint_matches_oracle(dev, shape):
    rng = np.random.default_rng(3)
    y = rng.uniform(-1, 1, shape + (1,)).astype(np.float32)
    w = O.laplacian_weights(0.1)
    want = O.heat_step_adjoint(y, w, 0.01, 1, 0.3)
    yd = T(y, dev)
    out = torch.empty_like(yd)
    call("pde_heat_step_adjoint_f32", yd.data_ptr(), out.data_ptr(), *shape, 1, 1, *w,
         np.float32(0.01), np.float32(0.3), N.stream_of(yd))
    assert same_bits(out.cpu().numpy(), want)


@pytest.mark.parametrize("first", [True, False])
@pytest.mark.parametrize("with_c2", [True, False])
@pytest.mark.parametrize("shape", [(24, 21, 1), (9, 10, 11)])
def test_wave_step_adjoint_matches_oracle(dev, first, with_c2, shape):
    rng = np.random.default_rng(4)
    w = O.laplacian_weights(0.1)
    gun, gvn, u2 = (rng.uniform(-1, 1, shape + (1,)).astype(np.float32) for _ in range(3))
    c2 = (1 + 0.3 * rng.uniform(-1, 1, shape + (1,))).astype(np.float32) if with_c2 else None
    gc2_0 = rng.uniform(-1, 1, shape + (1,)).astype(np.float32)
    gu2, gv, dc2 = O.wave_step_adjoint(gun, gvn, c2, u2, w, 0.7, 1, 0.05, first)
    d = {k: T(v, dev) for k, v in dict(gun=gun, gvn=gvn, u2=u2, gc2=gc2_0).items()}
    c2d = T(c2, dev) if with_c2 else None
    ogu2 = torch.empty_like(d["gun"])
    ogv = torch.empty_like(d["gun"])
    call("pde_wave_step_adjoint_f32", d["gun"].data_ptr(), d["gvn"].data_ptr(),
         c2d.data_ptr() if c2d is not None else None, d["u2"].data_ptr(), ogu2.data_ptr(),
         ogv.data_ptr(), d["gc2"].data_ptr(), *shape, 1, *w, np.float32(0.7), np.float32(0.05),
         1 if first else 0, N.stream_of(ogu2))
    assert same_bits(ogu2.cpu().numpy(), gu2)
    assert same_bits(ogv.cpu().numpy(), gv)
    assert same_bits(d["gc2"].cpu().numpy(), gc2_0 + dc2)


def _writes(ops):
    sign = {0: 0, 1: 1, 2: -1, 3: 1}
    w = [(int(o["dst"]), int(o["src"]), sign[int(o["kind"])]) for o in ops]
    return sorted(w, key=lambda t: (t[1], t[0]))


@pytest.mark.parametrize("legacy", [False, True])
def test_boundary_adjoint_matches_oracle_and_dot(dev, legacy):
    rng = np.random.default_rng(5)
    shape, C = (14, 13, 1), 2
    x = rng.uniform(-1, 1, shape + (C,)).astype(np.float32)
    xd = T(x, dev)
    if legacy:
        bc = LegacyGridBoundary(xd, 0.1, 1)
        bc.vonNeumann_BC("-X", 0.0)
        bc.impermeable("+Y")
    else:
        bc = GridBoundary(xd, 0.1, 1)
        bc.no_slip("ALL")
        bc.dirichlet_BC("+Y", 0.0, 0)
        bc.vonNeumann_BC("-X", 0.0, 1)
    out = bc(xd)              # all values 0: the boundary is linear here
    y = rng.uniform(-1, 1, shape + (C,)).astype(np.float32)
    yd = T(y, dev)
    gin = torch.empty_like(yd)
    bc.table.adjoint(yd, gin)
    want = O.boundary_linear_adjoint(y, _writes(bc.table.host_ops))
    assert same_bits(gin.cpu().numpy(), want)
    lhs = float(np.sum(out.cpu().numpy().astype(np.float64) * y))
    rhs = float(np.sum(x.astype(np.float64) * gin.cpu().numpy()))
    assert abs(lhs - rhs) <= 1e-5 * max(1.0, abs(lhs))


@pytest.mark.parametrize("shape,C", [((14, 13, 1), 2), ((70, 70, 70), 1)])
def test_boundary_adjoint_inplace_equals_out_of_place(dev, shape, C):
    """In-place transpose (one workgroup below 16384 sources, stage + store
    kernels above: the 70^3 case) equals the out-of-place one bit for bit."""
    print(f"bc adjoint in place {shape} C={C}", flush=True)
    rng = np.random.default_rng(8)
    x = T(rng.uniform(-1, 1, shape + (C,)).astype(np.float32), dev)
    bc = LegacyGridBoundary(x, 0.1, 1)
    bc.vonNeumann_BC("-X", 0.5)
    bc._ensure_table(x)
    y = T(rng.uniform(-1, 1, shape + (C,)).astype(np.float32), dev)
    want = torch.empty_like(y)
    bc.table.adjoint(y, want)
    got = y.clone()
    bc.table.adjoint_inplace(got)
    assert same_bits(got.cpu().numpy(), want.cpu().numpy())
    assert same_bits(want.cpu().numpy(), O.boundary_linear_adjoint(y.cpu().numpy(),
                                                                  _writes(bc.table.host_ops)))
