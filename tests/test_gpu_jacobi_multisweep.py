"""GPU checks of the multi-sweep Jacobi entry points through the C ABI, on
rectangular grids, ghost widths and sweep counts the projection cavity does
not use: k sweeps in one launch against k launches of the one-sweep kernels
(pde_jacobi_neumann0_f32 and its transpose), bit for bit, forward and
transposed (the rhs gradient accumulated onto a given array, or started from
0), and a batch against its instances."""
import numpy as np
import pytest
import torch

import pde_module_amd._native as N
from helpers import ghost_free_mask

pytestmark = pytest.mark.gpu

# (padded shape, ghost width, sweeps per launch)
CASES = [
    ((70, 200, 1), 1, 3),      # 64 x 64 regions, cut along both axes
    ((70, 200, 1), 2, 17),     # two ghost layers; 64 x 64 regions around 30-node outputs
    ((300, 45, 1), 1, 30),     # 128 x 128 regions (halo over 24); one piece along the columns
    ((9, 500, 1), 3, 5),       # three interior rows
    ((140, 140, 1), 1, 48),    # the largest halo: outputs of 32 nodes
    ((1, 90, 130), 2, 9),      # a 2D grid on the last two axes
]


def _rand(rng, shape, batch=None):
    lead = () if batch is None else (batch,)
    return rng.uniform(-1, 1, lead + shape + (1,)).astype(np.float32)


def _chain_forward(dev, p, rhs, shape, g, h2, inv, k):
    s = N.stream_of(p)
    a, b = p.clone(), torch.zeros_like(p)
    for _ in range(k):
        N.call("pde_jacobi_neumann0_f32", a.data_ptr(), rhs.data_ptr(), b.data_ptr(), *shape, g,
               h2, inv, s)
        a, b = b, a
    return a


def _chain_adjoint(dev, gin, grhs, shape, g, h2, inv, k):
    s = N.stream_of(gin)
    a, b = gin.clone(), torch.zeros_like(gin)
    for _ in range(k):
        N.call("pde_jacobi_neumann0_adjoint_f32", a.data_ptr(), b.data_ptr(), grhs.data_ptr(),
               *shape, g, h2, inv, s)
        a, b = b, a
    return a


@pytest.mark.parametrize("shape,g,k", CASES)
def test_multisweep_equals_one_sweep_launches(dev, shape, g, k):
    print(f"multisweep ABI {shape} ghost={g} sweeps={k}", flush=True)
    rng = np.random.default_rng(sum(shape) + 7 * g + k)
    h2, inv = np.float32(0.013), np.float32(0.25)
    mask = ghost_free_mask(shape, g, (1,))
    p = torch.from_numpy(_rand(rng, shape)).to(dev)
    rhs = torch.from_numpy(_rand(rng, shape)).to(dev)
    out = torch.zeros_like(p)
    N.call("pde_jacobi_neumann0_multisweep_f32", p.data_ptr(), rhs.data_ptr(), out.data_ptr(), 1,
           *shape, g, h2, inv, k, N.stream_of(p))
    want = _chain_forward(dev, p, rhs, shape, g, h2, inv, k)
    got_np, want_np = out.cpu().numpy(), want.cpu().numpy()
    assert np.array_equal(got_np[mask].view(np.int32), want_np[mask].view(np.int32))
    assert np.all(got_np[~mask] == 0)                 # ghosts are not written
    # transposed, accumulating onto a given rhs gradient
    gin = torch.from_numpy(_rand(rng, shape)).to(dev)
    g0 = _rand(rng, shape)
    gp, grhs = torch.zeros_like(gin), torch.from_numpy(g0.copy()).to(dev)
    N.call("pde_jacobi_neumann0_adjoint_multisweep_f32", gin.data_ptr(), None, gp.data_ptr(),
           grhs.data_ptr(), 1, *shape, g, h2, inv, k, 1, N.stream_of(gin))
    grhs_c = torch.from_numpy(g0.copy()).to(dev)
    gp_c = _chain_adjoint(dev, gin, grhs_c, shape, g, h2, inv, k)
    assert np.array_equal(gp.cpu().numpy().view(np.int32), gp_c.cpu().numpy().view(np.int32))
    assert np.array_equal(grhs.cpu().numpy().view(np.int32),
                          grhs_c.cpu().numpy().view(np.int32))
    # transposed, starting the rhs gradient from 0 (its ghosts are not written)
    junk = _rand(rng, shape)
    grhs0 = torch.from_numpy(junk.copy()).to(dev)
    N.call("pde_jacobi_neumann0_adjoint_multisweep_f32", gin.data_ptr(), None, gp.data_ptr(),
           grhs0.data_ptr(), 1, *shape, g, h2, inv, k, 0, N.stream_of(gin))
    zc = torch.zeros_like(gin)
    _chain_adjoint(dev, gin, zc, shape, g, h2, inv, k)
    got0, want0 = grhs0.cpu().numpy(), zc.cpu().numpy()
    assert np.array_equal(got0[mask].view(np.int32), want0[mask].view(np.int32))
    assert np.array_equal(got0[~mask].view(np.int32), junk[~mask].view(np.int32))
    # transposed from the sum of two inputs: the chain after pde_lincomb_f32 (1, 1)
    g2 = torch.from_numpy(_rand(rng, shape)).to(dev)
    gp2, grhs2 = torch.zeros_like(gin), torch.zeros_like(gin)
    N.call("pde_jacobi_neumann0_adjoint_multisweep_f32", gin.data_ptr(), g2.data_ptr(),
           gp2.data_ptr(), grhs2.data_ptr(), 1, *shape, g, h2, inv, k, 1, N.stream_of(gin))
    summed = torch.empty_like(gin)
    N.call("pde_lincomb_f32", gin.data_ptr(), g2.data_ptr(), None, None, np.float32(1.0),
           np.float32(1.0), np.float32(0), np.float32(0), 2, summed.data_ptr(), summed.numel(),
           N.stream_of(gin))
    grhs_s = torch.zeros_like(gin)
    gp_s = _chain_adjoint(dev, summed, grhs_s, shape, g, h2, inv, k)
    assert np.array_equal(gp2.cpu().numpy().view(np.int32), gp_s.cpu().numpy().view(np.int32))
    assert np.array_equal(grhs2.cpu().numpy().view(np.int32),
                          grhs_s.cpu().numpy().view(np.int32))


# (shape, ghost, sweeps, batch): a small batch, and one over 2^22 nodes (the
# forward in one LDS buffer of 64 x 64 nodes, the transpose in 128 x 128 regions)
@pytest.mark.parametrize("shape,g,k,B", [((70, 200, 1), 1, 7, 3), ((1026, 1026, 1), 1, 10, 4)])
def test_multisweep_batch_equals_instances(dev, shape, g, k, B):
    print(f"multisweep ABI batch of {B} {shape} sweeps={k}", flush=True)
    rng = np.random.default_rng(3)
    h2, inv = np.float32(0.013), np.float32(0.25)
    p = torch.from_numpy(_rand(rng, shape, B)).to(dev)
    rhs = torch.from_numpy(_rand(rng, shape, B)).to(dev)
    out = torch.zeros_like(p)
    N.call("pde_jacobi_neumann0_multisweep_f32", p.data_ptr(), rhs.data_ptr(), out.data_ptr(), B,
           *shape, g, h2, inv, k, N.stream_of(p))
    gin = torch.from_numpy(_rand(rng, shape, B)).to(dev)
    gp, grhs = torch.zeros_like(gin), torch.zeros_like(gin)
    N.call("pde_jacobi_neumann0_adjoint_multisweep_f32", gin.data_ptr(), None, gp.data_ptr(),
           grhs.data_ptr(), B, *shape, g, h2, inv, k, 1, N.stream_of(gin))
    for i in range(B):
        o1 = torch.zeros_like(p[i])
        N.call("pde_jacobi_neumann0_multisweep_f32", p[i].data_ptr(), rhs[i].data_ptr(),
               o1.data_ptr(), 1, *shape, g, h2, inv, k, N.stream_of(p))
        assert torch.equal(out[i].view(torch.int32), o1.view(torch.int32))
        g1, r1 = torch.zeros_like(gin[i]), torch.zeros_like(gin[i])
        N.call("pde_jacobi_neumann0_adjoint_multisweep_f32", gin[i].data_ptr(), None,
               g1.data_ptr(), r1.data_ptr(), 1, *shape, g, h2, inv, k, 1, N.stream_of(gin))
        assert torch.equal(gp[i].view(torch.int32), g1.view(torch.int32))
        assert torch.equal(grhs[i].view(torch.int32), r1.view(torch.int32))
