"""GPU parity of several heat steps per launch on LDS-resident tiles
(``pde_heat_multistep_f32``) against one step per launch, which the other
tests pin to the oracle: ``HeatStep.run`` on single grids (whole-grid tiles
and grids cut in both directions, legacy and functional tables with Neumann,
ramp and sine values, ghost widths 1 and 2, shifted groups, step counts that
leave a partial last launch) and ``HeatEnsemble`` with steps_per_launch > 1
(eager and graph-replayed) against stepping every instance alone: identical
bits everywhere; plus the 256^2 heat workload against the C oracle loop."""
import numpy as np
import pytest
import torch

from conftest import same_bits
from pde_module_amd import HeatStep
from pde_module_amd.batch import HeatEnsemble, PerInstanceHeat, TileConfig
from pde_module_amd.deprecated import GridBoundary as LegacyGridBoundary
from pde_module_amd.FDM import GridBoundary
from pde_module_amd.FDM.boundary import get_ramp_func, get_sine_func
from pde_module_amd.workloads import heat_ensemble_config

from test_multistep_plan import _functional, _legacy_mixed, _legacy_shift

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    return torch.equal(a.contiguous().view(torch.int32), b.contiguous().view(torch.int32))


CASES = [((40, 37), 1, _legacy_mixed, 8, 21), ((128, 128), 1, _functional, 4, 9),
         ((300, 140), 1, _legacy_mixed, 8, 16), ((133, 270), 2, _functional, 5, 13),
         ((260, 261), 1, _legacy_shift, 16, 40), ((129, 40), 1, _legacy_shift, 3, 7),
         ((258, 258), 1, lambda u, dx, g: LegacyGridBoundary(u, dx, g), 8, 24),
         ((600, 500), 1, _legacy_mixed, 8, 10)]


@pytest.mark.parametrize("shape,g,make,k,steps", CASES)
def test_heat_run_multistep_matches_single_steps(dev, shape, g, make, k, steps):
    print(f"multistep heat {shape} g={g} k={k} steps={steps}", flush=True)
    rng = np.random.default_rng(sum(shape) + k)
    u0 = torch.from_numpy(rng.uniform(-1, 1, shape + (1, 1)).astype(np.float32)).to(dev)
    dx, alpha = 0.01, 0.05
    dt = 0.2 * dx * dx / alpha
    multi = HeatStep(make(u0, dx, g), dx, g, alpha, dt)
    single = HeatStep(make(u0, dx, g), dx, g, alpha, dt)
    got = multi.run(u0.clone(), steps, t=0.3, steps_per_launch=k)
    assert multi._mplan is not None and multi._mplan[1] is not None, "multi-step path not taken"
    u = u0.clone()
    for _ in range(steps):
        u = single(u, 0.3)
    torch.cuda.synchronize()
    assert bits_equal(got, u)


def test_heat_run_graph_replay(dev):
    shape, g, k = (300, 140), 1, 8
    rng = np.random.default_rng(3)
    u0 = torch.from_numpy(rng.uniform(-1, 1, shape + (1, 1)).astype(np.float32)).to(dev)
    dx, alpha = 0.01, 0.05
    dt = 0.2 * dx * dx / alpha
    hs = HeatStep(_legacy_mixed(u0, dx, g), dx, g, alpha, dt)
    ref = HeatStep(_legacy_mixed(u0, dx, g), dx, g, alpha, dt)
    u = hs.run(u0.clone(), 16, steps_per_launch=k)      # buffers, plan uploaded
    torch.cuda.synchronize()
    start = hs._bufs[0]
    start.copy_(u0)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        out = hs.run(start, 16, steps_per_launch=k)   # 2 launches: back in start's pair
    assert out.data_ptr() == start.data_ptr()
    want = u0.clone()
    start.copy_(u0)
    for rep in range(3):
        graph.replay()
        torch.cuda.synchronize()
        for _ in range(16):
            want = ref(want)
        assert bits_equal(out, want), rep


@pytest.mark.parametrize("spl", [2, 8, 16])
def test_ensemble_multistep_matches_per_instance(dev, spl):
    from test_gpu_batch import _instances, _reference
    steps = 23
    print(f"ensemble multistep spl={spl}", flush=True)
    want = _reference(dev, steps)
    bcs, dxs, alphas, dts, u0s = _instances(dev)
    ens = HeatEnsemble(bcs, dxs, 1, alphas, dts, device=dev,
                       tiles=TileConfig(steps_per_launch=spl), graph_chunk=6)
    assert ens.multistep_ok(spl)
    ens.load(u0s)
    ens.run(steps, use_graph=False)
    for i, (got, w) in enumerate(zip(ens.fields(), want)):
        assert same_bits(got.cpu().numpy(), w), f"instance {i} {got.shape}"
    bcs, dxs, alphas, dts, u0s = _instances(dev)
    ens2 = HeatEnsemble(bcs, dxs, 1, alphas, dts, device=dev,
                        tiles=TileConfig(steps_per_launch=spl), graph_chunk=6)
    ens2.load(u0s)
    ens2.run(steps, use_graph=True)
    for got, w in zip(ens2.fields(), want):
        assert same_bits(got.cpu().numpy(), w)


def test_full_ensemble_workload_multistep(dev):
    """All 512 instances of the ensemble workload, 40 steps at 8 and 16 steps
    per launch (graph-replayed), bit-identical to one step per launch."""
    cfg = heat_ensemble_config(seed=3)
    fields = [torch.from_numpy(cfg.u0(i, 3)).to(dev) for i in range(len(cfg.n))]
    mk = lambda: [LegacyGridBoundary(f, float(cfg.dx[i]), 1) for i, f in enumerate(fields)]
    args = ([float(d) for d in cfg.dx], 1, [float(a) for a in cfg.alpha],
            [float(d) for d in cfg.dt])
    base = HeatEnsemble(mk(), *args, device=dev)
    base.load(fields)
    base.run(40)
    for spl in (8, 16):
        print(f"full ensemble multistep spl={spl}", flush=True)
        ens = HeatEnsemble(mk(), *args, device=dev, tiles=TileConfig(steps_per_launch=spl))
        ens.load(fields)
        ens.run(40)
        torch.cuda.synchronize()
        for i, (a, b) in enumerate(zip(ens.fields(), base.fields())):
            assert bits_equal(a, b), f"instance {i} n={cfg.n[i]}"


def test_heat_workload_multistep_matches_oracle(dev, corc):
    """The 256^2 heat workload (Dirichlet 0, 8 Gaussian bumps) for 64 steps at
    16 steps per launch against the C oracle loop, bit for bit."""
    from oracle import np_oracle as O
    from pde_module_amd.workloads import heat_2d_config
    cfg = heat_2d_config(seed=0)
    u0 = cfg.u0 if hasattr(cfg, "u0") else cfg["u0"]
    u0 = np.asarray(u0, np.float32)
    ud = torch.from_numpy(u0).to(dev)
    dx, alpha, dt = (float(getattr(cfg, k)) for k in ("dx", "alpha", "dt"))
    bc = LegacyGridBoundary(ud, dx, 1)
    hs = HeatStep(bc, dx, 1, alpha, dt)
    got = hs.run(ud.clone(), 64, steps_per_launch=16).cpu().numpy()
    u = u0.copy()
    w = O.laplacian_weights(dx)           # from the FP64 dx, as the module's
    for _ in range(64):
        u2 = corc.bc_legacy(u, bc.boundary_ijk_indices, bc.boundary_interior, bc.boundary_type,
                            bc.boundary_value, dx)
        u = corc.forward_euler(u2, corc.laplacian(u2, w, np.float32(alpha), 1), np.float32(dt))
    m = np.zeros(u.shape, bool)
    m[1:-1, 1:-1] = True
    assert same_bits(got, u, m)
