"""GPU parity of the one-launch heat step (``pde_heat_step_ring_f32``, the
boundary folded into the step's launch) against the two-launch form (in-place
boundary, then ``pde_heat_step_f32``), which the other tests pin to the C
restatement: every returned state and every boundary-applied input identical
bit for bit over several steps, for legacy and functional tables (constant,
Neumann and time-ramped values), ghost widths 1 and 2, grids from one tile
to many, and under graph replay."""
import numpy as np
import pytest
import torch

from pde_module_amd import HeatStep
from pde_module_amd.deprecated import GridBoundary as LegacyGridBoundary
from pde_module_amd.FDM import GridBoundary
from pde_module_amd.FDM.boundary import get_ramp_func

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    return torch.equal(a.contiguous().view(torch.int32), b.contiguous().view(torch.int32))


def _legacy(u, dx, g):
    return LegacyGridBoundary(u, dx, g)


def _mixed(u, dx, g):
    bc = LegacyGridBoundary(u, dx, g)
    bc.vonNeumann_BC("-X", 0.5)
    bc.dirichlet_BC("+Y", 0.75)
    return bc


def _ramp(u, dx, g):
    bc = GridBoundary(u, dx, g)
    bc.dirichlet_BC("ALL", get_ramp_func(3.0, 0.4))
    bc.vonNeumann_BC("+Y", 0.25)
    return bc


CASES = [((258, 258), 1, _legacy), ((40, 37), 1, _mixed), ((9, 40), 1, _legacy),
         ((300, 77), 1, _ramp), ((61, 525), 2, _mixed), ((600, 500), 1, _legacy)]


@pytest.mark.parametrize("shape,g,make", CASES)
def test_one_launch_heat_step_matches_two_launches(dev, shape, g, make):
    print(f"one-launch heat step {shape} g={g} {make.__name__}", flush=True)
    rng = np.random.default_rng(sum(shape))
    u0 = torch.from_numpy(rng.uniform(-1, 1, shape + (1, 1)).astype(np.float32)).to(dev)
    dx, alpha = 0.01, 0.05
    dt = 0.2 * dx * dx / alpha
    one = HeatStep(make(u0, dx, g), dx, g, alpha, dt)
    two = HeatStep(make(u0, dx, g), dx, g, alpha, dt, one_launch=False)
    ua, ub = u0.clone(), u0.clone()
    for n in range(6):
        t = 0.1 * n
        na, nb = one(ua, t), two(ub, t)
        torch.cuda.synchronize()
        assert one._ring[1] is not None, "the one-launch path was not taken"
        assert bits_equal(na, nb), f"step {n}: next state differs"
        assert bits_equal(one._bufs[0] if na is one._bufs[1] else one._bufs[1],
                          two._bufs[0] if nb is two._bufs[1] else two._bufs[1]), \
            f"step {n}: boundary-applied input differs"
        ua, ub = na, nb


def test_one_launch_heat_step_graph_replay(dev):
    shape, g = (130, 300), 1
    rng = np.random.default_rng(7)
    u0 = torch.from_numpy(rng.uniform(-1, 1, shape + (1, 1)).astype(np.float32)).to(dev)
    dx, alpha = 0.01, 0.05
    dt = 0.2 * dx * dx / alpha
    one = HeatStep(_mixed(u0, dx, g), dx, g, alpha, dt)
    two = HeatStep(_mixed(u0, dx, g), dx, g, alpha, dt, one_launch=False)
    u = one(one(u0))
    torch.cuda.synchronize()
    a = one._bufs[0]
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        one(one(a))
    a.copy_(u0)
    ub = u0.clone()
    for rep in range(3):
        graph.replay()
        torch.cuda.synchronize()
        ub = two(two(ub))
        torch.cuda.synchronize()
        assert bits_equal(a, ub), f"replay {rep}"


def test_heat_rollout_forward_matches_two_launch_loop(dev):
    """HeatRollout's forward (one launch per step) against the two-launch
    HeatStep loop, bit for bit, with checkpoints every 7 steps."""
    from pde_module_amd.rollout import HeatRollout
    shape, g, steps = (70, 45), 1, 30
    rng = np.random.default_rng(11)
    u0 = torch.from_numpy(rng.uniform(-1, 1, shape + (1, 1)).astype(np.float32)).to(dev)
    dx, alpha = 0.01, 0.05
    dt = 0.2 * dx * dx / alpha
    ro = HeatRollout(_mixed(u0, dx, g), dx, g, alpha, dt, steps, checkpoint_every=7)
    uN, ckpt = ro.forward_states(u0)
    assert ro._hs._ring[1] is not None, "the one-launch path was not taken"
    two = HeatStep(_mixed(u0, dx, g), dx, g, alpha, dt, one_launch=False)
    ub = u0.clone()
    for n in range(steps):
        ub = two(ub)
        if n + 1 in ckpt:
            assert bits_equal(ckpt[n + 1], ub), f"checkpoint {n + 1}"
    torch.cuda.synchronize()
    assert bits_equal(uN, ub)


def test_one_launch_heat_step_reads_a_device_time_tensor(dev):
    """The ramp's time as a 1-element FP32 device tensor (the form a captured
    loop updates in place) gives the same states as the two launches."""
    shape, g = (90, 130), 1
    rng = np.random.default_rng(13)
    u0 = torch.from_numpy(rng.uniform(-1, 1, shape + (1, 1)).astype(np.float32)).to(dev)
    dx, alpha = 0.01, 0.05
    dt = 0.2 * dx * dx / alpha
    one = HeatStep(_ramp(u0, dx, g), dx, g, alpha, dt)
    two = HeatStep(_ramp(u0, dx, g), dx, g, alpha, dt, one_launch=False)
    t = torch.zeros(1, dtype=torch.float32, device=dev)
    ua, ub = u0.clone(), u0.clone()
    for n in range(5):
        t.fill_(0.7 * n)
        ua, ub = one(ua, t), two(ub, t)
        torch.cuda.synchronize()
        assert one._ring[1] is not None
        assert bits_equal(ua, ub), f"step {n}"
