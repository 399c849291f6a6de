"""GPU parity of the batched heat ensemble: every instance bit-identical to
stepping it alone with HeatStep, for every tile configuration and both launch
modes, eager and graph-replayed; anchored to the C FP32 oracle on a few
instances; autotuning leaves the state untouched."""
import numpy as np
import pytest
import torch

from conftest import same_bits
from pde_module_amd.batch import HeatEnsemble, PerInstanceHeat, TileConfig
from pde_module_amd.deprecated import GridBoundary as LegacyGridBoundary
from pde_module_amd.FDM import GridBoundary
from pde_module_amd.workloads import heat_ensemble_config

pytestmark = pytest.mark.gpu

SIDES = [(3, 3), (5, 9), (18, 66), (66, 18), (70, 70), (130, 131), (260, 35), (33, 515),
         (1, 1), (258, 258)]


def _instances(dev, seed=0):
    """Mixed legacy / functional boundaries, uneven sizes, nonzero values."""
    rng = np.random.default_rng(seed)
    bcs, dxs, alphas, dts, u0s = [], [], [], [], []
    for i, (a, b) in enumerate(SIDES):
        if (a, b) == (1, 1):
            continue
        shape = (a + 2, b + 2, 1, 1)
        u0 = torch.from_numpy(rng.uniform(-1, 1, shape).astype(np.float32)).to(dev)
        dx = 1.0 / max(a, b, 2)
        if i % 3 == 0:
            bc = LegacyGridBoundary(u0, dx, 1)
            bc.dirichlet_BC("ALL", 0.0)
        elif i % 3 == 1:
            bc = GridBoundary(u0, dx, 1)
            bc.dirichlet_BC("ALL", 0.25)
            bc.vonNeumann_BC("+X", -0.5)
        else:
            bc = LegacyGridBoundary(u0, dx, 1)
            bc.dirichlet_BC("ALL", 1.0)
            bc.vonNeumann_BC("-Y", 0.3)
        alpha = 0.01 + 0.01 * i
        bcs.append(bc)
        dxs.append(dx)
        alphas.append(alpha)
        dts.append(0.2 * dx * dx / alpha)
        u0s.append(u0)
    return bcs, dxs, alphas, dts, u0s


def _reference(dev, steps):
    bcs, dxs, alphas, dts, u0s = _instances(dev)
    ref = PerInstanceHeat(bcs, dxs, 1, alphas, dts)
    ref.load(u0s)
    ref.run(steps)
    return [f.cpu().numpy() for f in ref.fields()]


@pytest.mark.parametrize("rows", [1, 2, 4, 8])
@pytest.mark.parametrize("waves", [1, 2, 4])
@pytest.mark.parametrize("blocks,xcd", [(0, 8), (7, 8), (0, 1), (16, 1)])
def test_ensemble_matches_per_instance(dev, rows, waves, blocks, xcd):
    steps = 13
    print(f"ensemble rows={rows} waves={waves} blocks={blocks} xcd={xcd}", flush=True)
    want = _reference(dev, steps)
    bcs, dxs, alphas, dts, u0s = _instances(dev)
    ens = HeatEnsemble(bcs, dxs, 1, alphas, dts, device=dev,
                       tiles=TileConfig(rows, waves, blocks, xcd))
    ens.load(u0s)
    ens.run(steps, use_graph=False)
    for i, (got, w) in enumerate(zip(ens.fields(), want)):
        assert same_bits(got.cpu().numpy(), w), f"instance {i} {got.shape}"


def test_ensemble_graph_matches_eager_and_oracle(dev, corc):
    steps = 23
    want = _reference(dev, steps)
    bcs, dxs, alphas, dts, u0s = _instances(dev)
    ens = HeatEnsemble(bcs, dxs, 1, alphas, dts, device=dev, graph_chunk=6)
    ens.load(u0s)
    ens.run(steps, use_graph=True)      # 3 replays of a 6-step graph + 5 eager steps
    for got, w in zip(ens.fields(), want):
        assert same_bits(got.cpu().numpy(), w)
    # a second run continues from the state (graph re-captured for the swapped buffers)
    ens.run(7, use_graph=True)
    want2 = _reference(dev, steps + 7)
    for got, w in zip(ens.fields(), want2):
        assert same_bits(got.cpu().numpy(), w)
    # anchor two legacy instances to the C oracle loop
    bcs, dxs, alphas, dts, u0s = _instances(dev)
    for i in (0, 3):
        bc, u = bcs[i], u0s[i].cpu().numpy()
        from oracle import np_oracle as O
        w = O.laplacian_weights(dxs[i])
        for _ in range(steps):
            u2 = corc.bc_legacy(u, bc.boundary_ijk_indices, bc.boundary_interior,
                                bc.boundary_type, bc.boundary_value, dxs[i])
            L = corc.laplacian(u2, w, np.float32(alphas[i]), 1)
            u = corc.forward_euler(u2, L, np.float32(dts[i]))
        M = np.zeros(u.shape, bool)
        M[1:-1, 1:-1] = True
        assert same_bits(want[i], u, M)


def test_autotune_restores_state(dev):
    bcs, dxs, alphas, dts, u0s = _instances(dev)
    ens = HeatEnsemble(bcs, dxs, 1, alphas, dts, device=dev)
    ens.load(u0s)
    before = [f.clone() for f in ens.fields()]
    cfg = ens.autotune(candidates=[TileConfig(4, 4, 0), TileConfig(2, 2, 16)], steps=2, reps=1)
    assert isinstance(cfg, TileConfig) and len(ens.timings) == 2
    for a, b in zip(before, ens.fields()):
        assert torch.equal(a, b)
    ens.run(11)
    for got, w in zip(ens.fields(), _reference(dev, 11)):
        assert same_bits(got.cpu().numpy(), w)


def test_ensemble_workload_subset_matches_per_instance(dev):
    """Eight instances of the 512-grid workload (sizes 64..1024), 5 steps."""
    cfg = heat_ensemble_config(seed=3)
    idx = np.argsort(cfg.n)[::64][:8]
    fields = [torch.from_numpy(cfg.u0(int(i), 3)).to(dev) for i in idx]
    mk = lambda: [LegacyGridBoundary(f, float(cfg.dx[i]), 1) for f, i in zip(fields, idx)]
    args = ([float(cfg.dx[i]) for i in idx], 1, [float(cfg.alpha[i]) for i in idx],
            [float(cfg.dt[i]) for i in idx])
    ens = HeatEnsemble(mk(), args[0], 1, args[2], args[3], device=dev)
    ens.load(fields)
    ens.run(5)
    ref = PerInstanceHeat(mk(), *args)
    ref.load(fields)
    ref.run(5)
    for a, b in zip(ens.fields(), ref.fields()):
        assert same_bits(a.cpu().numpy(), b.cpu().numpy())


def test_full_ensemble_workload_matches_per_instance_and_oracle(dev, corc):
    """All 512 instances of the ensemble workload (sides 64..1024 log-uniform,
    seed 3, Dirichlet-0 walls), 20 steps through the graph-replayed batch:
    every instance bit-identical to stepping it alone, and six instances
    spanning the size range bit-identical to the C oracle's loop
    (examples/2D_Heat_Equation.py:58-62: boundary, Laplacian, ForwardEuler)."""
    from oracle import np_oracle as O
    steps = 20
    cfg = heat_ensemble_config(seed=3)
    print(f"full ensemble: {len(cfg.n)} instances, {steps} steps", flush=True)
    u0s = [cfg.u0(i, 3) for i in range(len(cfg.n))]
    fields = [torch.from_numpy(u).to(dev) for u in u0s]
    mk = lambda: [LegacyGridBoundary(f, float(cfg.dx[i]), 1) for i, f in enumerate(fields)]
    dxs = [float(d) for d in cfg.dx]
    alphas = [float(a) for a in cfg.alpha]
    dts = [float(d) for d in cfg.dt]
    ens = HeatEnsemble(mk(), dxs, 1, alphas, dts, device=dev, graph_chunk=10)
    ens.load(fields)
    ens.run(steps, use_graph=True)
    ref = PerInstanceHeat(mk(), dxs, 1, alphas, dts)
    ref.load(fields)
    ref.run(steps)
    torch.cuda.synchronize()
    for i, (a, b) in enumerate(zip(ens.fields(), ref.fields())):
        assert torch.equal(a.view(torch.int32), b.view(torch.int32)), f"instance {i} n={cfg.n[i]}"
    order = np.argsort(cfg.n)
    picks = order[np.linspace(0, len(order) - 1, 6).round().astype(int)]
    got = ens.fields()
    for i in picks:
        bc = LegacyGridBoundary(fields[i], dxs[i], 1)
        u = u0s[i].copy()
        w = O.laplacian_weights(dxs[i])
        for _ in range(steps):
            u2 = corc.bc_legacy(u, bc.boundary_ijk_indices, bc.boundary_interior,
                                bc.boundary_type, bc.boundary_value, dxs[i])
            L = corc.laplacian(u2, w, np.float32(alphas[i]), 1)
            u = corc.forward_euler(u2, L, np.float32(dts[i]))
        M = np.zeros(u.shape, bool)
        M[1:-1, 1:-1] = True
        print(f"  oracle anchor: instance {i} n={cfg.n[i]}", flush=True)
        assert same_bits(got[i].cpu().numpy(), u, M), f"instance {i} n={cfg.n[i]}"
