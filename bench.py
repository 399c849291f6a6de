"""Benchmark: 3D 7-point FP32 Laplacian on a 512^3 periodic grid (514^3 padded,
one ghost layer), one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

A step is one periodic Laplacian application over the whole grid: one launch
of the wrapped-index Laplacian (PeriodicLaplacian), which gives bit for bit the
values of a periodic ghost fill followed by the drop-in Laplacian module.  The
drop-in path itself (PeriodicBoundary + Laplacian, two modules, two launches)
is timed beside it.  The inputs (s ~ U[-1, 1], seeded) are resident in HBM
before timing starts.  Printed: one JSON line with GLUPS (interior lattice
updates per second / 1e9), the step kernel's achieved bandwidth against the
8 TB/s HBM peak (HIP events on the launch stream, algorithmic 8 bytes per
update), the drop-in Laplacian kernel's, the Grad, Div and fused lap+grad+div
passes on the same grid, the 2048^2 wave workload's forward+adjoint vs forward
time, and the CPU restatement timed on a bounded slab of the same workload on
this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
N = 512
G = 1
P = N + 2 * G
LUPS = N ** 3
LAP_BYTES = 8          # read s, write lap
GRAD_BYTES = 16        # read s, write vec3
DIV_BYTES = 16         # read vec3, write s
FUSED_BYTES = 36       # read s + vec3, write lap + vec3 + div
# two reversed transposed wave steps: read u, v, g_u, g_v, c2, gc2; write
# u, v, g_u, g_v, gc2 (per padded node)
WAVE_PAIR_BYTES = 44


PREWARM_S = 0.3     # untimed seconds of the step before the warm-up steps
MULTI_SPL = 12      # heat steps per launch on LDS-resident tiles (256^2 workload)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    # the first ~10 ms of back-to-back launches run slower (clock ramp): warm
    # up past that
    p.add_argument("--warmup", type=int, default=50)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--cpu-seconds", type=float, default=12.0)
    p.add_argument("--quick", action="store_true",
                   help="operators only: skip the wave and ensemble workloads (profiling runs)")
    p.add_argument("--only-wave", action="store_true",
                   help="profiling runs: only the 2048^2 wave workload, its JSON on stdout")
    p.add_argument("--wave-steps", type=int, default=2000,
                   help="steps of the wave workload (profiling runs shorten it)")
    p.add_argument("--only-heat", action="store_true",
                   help="profiling runs: only the heat workloads (512-grid ensemble, 256^2 "
                        "loop), their JSON on stdout")
    p.add_argument("--only-ldc", action="store_true",
                   help="only the module-path lid-driven cavity (the reference's benchmark harness)")
    p.add_argument("--only-cavity", action="store_true",
                   help="profiling runs: only the 1024^2 projection cavity, its JSON on stdout")
    p.add_argument("--cavity-steps", type=int, default=2000,
                   help="steps of the single-cavity rollouts (profiling runs shorten it)")
    return p.parse_args()


def progress(msg):
    """One progress line on stderr (the JSON result is the only stdout line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def pmc_traffic():
    """HBM bytes per launch of the 3D Laplacian kernels from the committed
    rocprofv3 counter summary (FETCH_SIZE + WRITE_SIZE, scaled by the
    copy-kernel calibration): {kernel: bytes}, or None when no summary is
    present."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "lap3d_512_pmc.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return {k: float(v) for k, v in d["bytes_per_launch"].items()}
    except (OSError, KeyError, ValueError, AttributeError):
        return None


def host_cpu_model():
    """The host CPU's model name (/proc/cpuinfo) and the cores this process may
    run on (sched_getaffinity)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "affinity_cores": len(os.sched_getaffinity(0))}


def cpu_baseline(seconds):
    """The C FP32 restatement (single thread, the serial reference-equivalent
    CPU path) on slabs of the 512^3 workload: 512 x 512 x 512 planes of the
    periodic grid are too slow to repeat, so slabs of 32 interior planes (8.4M
    updates) are timed until about `seconds` of CPU work."""
    from oracle import c_oracle
    from oracle.np_oracle import laplacian_weights
    rng = np.random.default_rng(0)
    slab = rng.uniform(-1, 1, (32 + 2, P, P, 1)).astype(np.float32)
    out = np.empty_like(slab)
    w = laplacian_weights(1.0 / N)
    lib = c_oracle.lib()
    res = {}
    for threads in (1, c_oracle.max_threads()):
        reps, t0 = 0, time.perf_counter()
        while True:
            lib.orc_laplacian_f32(slab, out, slab.shape[0], P, P, 1, G, w[0], w[1], w[2],
                                  np.float32(1.0), threads)
            reps += 1
            el = time.perf_counter() - t0
            if el > seconds / 2 or (threads == 1 and reps >= 40):
                break
        res[threads] = reps * 32 * N * N / el / 1e9
    return res


def wave_cpu_timing(seconds):
    """The 2048^2 wave workload's forward step on the host: the C FP32
    restatement of the boundary (legacy Dirichlet 0 on every face) and of the
    u/v update (oracle/pde_oracle.c, the reference loop's arithmetic), single
    thread and all threads, on the same grid and c^2, from the seeded initial
    state (the source injection, a few nodes, is left out).  Steps are timed
    until about `seconds` of CPU work per thread count."""
    from oracle import c_oracle
    from oracle import np_oracle as O
    from pde_module_amd.workloads import wave_2d_config
    cfg = wave_2d_config(seed=0, steps=1)
    c2 = np.ascontiguousarray(cfg.c2, dtype=np.float32)
    rng = np.random.default_rng(0)
    u = rng.uniform(-1e-3, 1e-3, c2.shape).astype(np.float32)
    v = np.zeros_like(u)
    import torch
    from pde_module_amd.deprecated import GridBoundary as LegacyGridBoundary
    bc = LegacyGridBoundary(torch.zeros(c2.shape), cfg.dx, cfg.g)     # default Dirichlet 0
    tabs = (bc.boundary_ijk_indices, bc.boundary_interior, bc.boundary_type, bc.boundary_value)
    w = O.laplacian_weights(cfg.dx)
    interior = (cfg.n) ** 2
    res = {"grid": [cfg.n, cfg.n], "note": "C FP32 restatement (boundary + u/v update), the "
           "source injection left out"}
    for threads in (1, c_oracle.max_threads()):
        uu, vv, reps, t0 = u.copy(), v.copy(), 0, time.perf_counter()
        while True:
            u2 = c_oracle.bc_legacy(uu, *tabs, cfg.dx)
            uu, vv = c_oracle.wave_step_core(u2, vv, c2, w, 1.0, cfg.g, float(cfg.dt), reps == 0,
                                             nthreads=threads)
            reps += 1
            el = time.perf_counter() - t0
            if el > seconds / 2 or reps >= 200:
                break
        res[f"threads_{threads}"] = {"ms_per_step": round(el / reps * 1e3, 3),
                                     "glups": round(reps * interior / el / 1e9, 4),
                                     "steps_timed": reps}
    return res


def wave_adjoint_timing(dev, steps=2000, K=50, rec_distance=None, variants=True):
    """2D acoustic workload (2048^2, smoothed random c, Ricker source at the
    centre, 64 receivers, Dirichlet-0 walls, 2000 steps): time of the plain
    forward loop vs forward + adjoint with checkpoints every K steps, all as
    HIP-graph-replayed segments.  The checkpointed gradient walks each segment
    back from its end (WaveRollout(reverse=True)); the recompute form and the
    store-every-state form are timed beside it (variants=True).  J = sum of
    squared receiver misfit against data from the unperturbed model (c = 1).

    rec_distance=None puts the receivers on the wall row, which the front does
    not reach in 2000 steps (J = 0); with a distance the wave arrives, and the
    reversed gradient is checked against the recompute path's (records and
    final state bit-identical, dJ/dc2 within 2e-5 relative)."""
    import torch
    from pde_module_amd.deprecated import GridBoundary as LegacyGridBoundary
    from pde_module_amd.rollout import WaveRollout
    from pde_module_amd.workloads import wave_2d_config
    cfg = wave_2d_config(seed=0, steps=steps, rec_distance=rec_distance)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    c2 = T(cfg.c2)
    u0 = torch.zeros_like(c2)
    v0 = torch.zeros_like(c2)
    bc = LegacyGridBoundary(u0, cfg.dx, cfg.g)
    mk = lambda k, rev=False: WaveRollout(bc, cfg.dx, cfg.g, float(cfg.dt), steps,
                                          checkpoint_every=k, src_idx=T(cfg.src_idx),
                                          src_wavelet=T(cfg.wavelet), rec_idx=T(cfg.rec_idx),
                                          use_graph=True, reverse=rev)
    plain, ro, ro_rev = mk(steps), mk(K), mk(K, True)
    _, _, d = plain.forward_states(torch.ones_like(c2), u0, v0)
    d = d.clone()

    def timed(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        return best

    t_fwd = timed(lambda: plain.forward_states(c2, u0, v0))
    out = {}

    def grad_of(rollout, key):
        def run():
            c = c2.clone().requires_grad_(True)
            uN, _, r = rollout(c, u0, v0)
            J = ((r - d) ** 2).sum()
            J.backward()
            out[key] = (J.detach(), c.grad, r.detach(), uN.detach())
        return run

    t_rev = timed(grad_of(ro_rev, "rev"))
    res = {"grid": [cfg.n, cfg.n], "steps": steps, "checkpoint_every": K,
           "receiver_row_from_source": int(cfg.rec_idx[0] % c2.shape[1]) - (cfg.g + cfg.n // 2),
           "fwd_ms_per_step": round(t_fwd / steps * 1e3, 5),
           "fwd_adj_ms_per_step": round(t_rev / steps * 1e3, 5),
           "ratio": round(t_rev / t_fwd, 3),
           "J": float(out["rev"][0]), "max_abs_dJ_dc2": float(out["rev"][1].abs().max()),
           "mode": "checkpoint every K steps; the backward walks each segment back from the "
                   "checkpoint at its end, two transposed steps per launch, rebuilding the "
                   "states (boundary-overwritten scalars recorded by the forward) instead of "
                   "recomputing the segment; steps 0 and 1 (plain Euler first step) are "
                   "recomputed"}
    if variants:
        t_fa = timed(grad_of(ro, "rec"))
        ro_all = mk(None)
        t_all = timed(grad_of(ro_all, "all"), reps=2)
        del ro_all
        res["recompute"] = {"fwd_adj_ms_per_step": round(t_fa / steps * 1e3, 5),
                            "ratio": round(t_fa / t_fwd, 3)}
        res["store_all"] = {"fwd_adj_ms_per_step": round(t_all / steps * 1e3, 5),
                            "ratio": round(t_all / t_fwd, 3),
                            "history_GB": round((steps + 1) * c2.numel() * 4 / 1e9, 1)}
        res["note"] = ("graph-replayed 50-step segments; recompute: each segment recomputed "
                       "from its checkpoint in the backward; store_all: every state kept")
    else:
        grad_of(ro, "rec")()
    if rec_distance is not None:
        bits = lambda a, b: torch.equal(a.contiguous().view(torch.int32),
                                        b.contiguous().view(torch.int32))
        J, g, r, uN = out["rev"]
        Jr, gr, rr, uNr = out["rec"]
        e = float((g.double() - gr.double()).norm() / gr.double().norm())
        same = bits(r, rr) and bits(uN, uNr)
        res["check"] = {"J_gt_0": float(J) > 0, "records_and_uN_bitwise_equal": same,
                        "dJ_dc2_rel_vs_recompute": e}
        if not (float(J) > 0 and same and e < 2e-5):
            raise AssertionError(f"wave gradient check failed: {res['check']}")
        res["data_note"] = (f"receiver row {rec_distance} cells from the source (the wave "
                            "arrives); everything else as the workload")

    # the pair kernel alone: one reversed segment's K/2 launches (steps
    # 2K-1 .. K, which need at least 2K recorded steps), HIP events on the
    # launch stream (eager launches; state from the last backward)
    if steps < 2 * K:
        raise ValueError(f"the wave workload needs at least {2 * K} steps")
    ro_rev._lu.zero_()
    ro_rev._lv.zero_()
    ro_rev._set_counter(2 * K - 1)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ro_rev._rev_segment(K, ro_rev._c2, True)
    torch.cuda.synchronize()
    ro_rev._set_counter(2 * K - 1)
    e0.record()
    ro_rev._rev_segment(K, ro_rev._c2, True)
    e1.record()
    torch.cuda.synchronize()
    pair_us = e0.elapsed_time(e1) * 1e3 / (K // 2)
    pair_bytes = WAVE_PAIR_BYTES * c2.numel()
    res["pair_kernel"] = {"kernel": "wave_pair_kernel", "avg_launch_us": round(pair_us, 2),
                          "algorithmic_bytes_per_launch": pair_bytes,
                          "GBs": round(pair_bytes / pair_us / 1e3, 1),
                          "frac": round(pair_bytes / pair_us / 1e3 / HBM_PEAK_GBS, 4),
                          "note": "algorithmic bytes / time against 8 TB/s; the launch's "
                                  "fields (~170 MB) partly stay in the 256 MB Infinity "
                                  "Cache"}
    return res


def heat_ensemble_timing(dev, steps=500, baseline_steps=10, seed=3):
    """Batched heat workload: 512 independent 2D grids, sides log-uniform in
    [64, 1024] (seed 3: 89.3 M interior points), diffusivity U[0.005, 0.05],
    dt = 0.2 dx^2 / alpha, Dirichlet 0 walls (the legacy GridBoundary
    default), 8 Gaussian bumps each.  The ensemble (autotuned tiles,
    graph-replayed 50-step chunks) runs all 500 steps; the per-instance
    baseline (one HeatStep per grid, launched grid by grid) is timed over
    ``baseline_steps`` steps of every grid."""
    import torch
    from pde_module_amd.batch import HeatEnsemble, PerInstanceHeat, TileConfig
    from pde_module_amd.deprecated import GridBoundary as LegacyGridBoundary
    from pde_module_amd import _native as NV
    from pde_module_amd.workloads import heat_ensemble_config
    cfg = heat_ensemble_config(seed=seed)
    count = len(cfg.n)
    fields = [torch.from_numpy(cfg.u0(i, seed)).to(dev) for i in range(count)]
    mk = lambda: [LegacyGridBoundary(f, float(cfg.dx[i]), cfg.g) for i, f in enumerate(fields)]
    dxs, alphas, dts = (list(map(float, a)) for a in (cfg.dx, cfg.alpha, cfg.dt))
    ens = HeatEnsemble(mk(), dxs, cfg.g, alphas, dts, device=dev)
    ens.load(fields)
    best = ens.autotune(steps=6, reps=2)
    ens.load(fields)
    ens.run(2 * ens.graph_chunk)               # capture + warm up
    torch.cuda.synchronize()
    ens.load(fields)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.run(steps)
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) / steps
    # kernel-only time per step (HIP events on the launch stream): the best
    # one-step-per-launch tile config's step kernel, and the multi-step launch
    # divided by its steps when that is the tuned config
    one = min((k for k in ens.timings if k[4] == 1), key=lambda k: ens.timings[k])
    one_cfg = TileConfig(*one)
    tiles, nt = ens._tiles(one_cfg)
    a, b = ens.bufs
    stream = torch.cuda.current_stream(dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 20
    e0.record(stream)
    for _ in range(reps):
        NV.call("pde_heat_batch_step_f32", a.data_ptr(), b.data_ptr(), ens.grid_dev.data_ptr(),
                ens.n, tiles.data_ptr(), nt, one_cfg.rows_per_lane, one_cfg.waves,
                one_cfg.blocks, NV.stream_of(a))
    e1.record(stream)
    e1.synchronize()
    t_kernel = e0.elapsed_time(e1) / reps * 1e-3
    # a plain copy of the packed buffer: the bandwidth yardstick, and the
    # calibration of counter runs over this workload
    e0.record(stream)
    for _ in range(reps):
        NV.call("pde_copy_f32", a.data_ptr(), b.data_ptr(), a.numel(), NV.stream_of(a))
    e1.record(stream)
    e1.synchronize()
    t_copy = e0.elapsed_time(e1) / reps * 1e-3
    t_multi = None
    if best.steps_per_launch > 1:
        spl = best.steps_per_launch
        e0.record(stream)
        for _ in range(reps):
            ens._launch_multistep(a, b, spl, spl)
        e1.record(stream)
        e1.synchronize()
        t_multi = e0.elapsed_time(e1) / reps * 1e-3
    padded = int((ens.p1 * ens.p2).sum())
    ref = PerInstanceHeat(mk(), dxs, cfg.g, alphas, dts)
    ref.load(fields)
    ref.run(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ref.run(baseline_steps)
    torch.cuda.synchronize()
    t_ref = (time.perf_counter() - t0) / baseline_steps
    return {"instances": count, "interior_points": ens.points, "steps": steps,
            "batched_ms_per_step": round(t_batch * 1e3, 4),
            "per_instance_ms_per_step": round(t_ref * 1e3, 4),
            "speedup": round(t_ref / t_batch, 2),
            "batched_glups": round(ens.points / t_batch / 1e9, 1),
            "tile_config": ({"steps_per_launch": best.steps_per_launch}
                            if best.steps_per_launch > 1 else
                            {"rows_per_lane": best.rows_per_lane, "waves": best.waves,
                             "blocks": best.blocks, "xcd_slots": best.xcd_slots}),
            "autotune_ms_per_step": {(f"{k[4]} steps/launch" if k[4] > 1 else
                                      f"{k[0]}x{k[1]}/{k[2]}/{k[3]}"): round(v, 4)
                                     for k, v in sorted(ens.timings.items())},
            "multistep_kernel": (None if t_multi is None else {
                "kernel": "heat_mstep_kernel", "steps_per_launch": best.steps_per_launch,
                "us_per_launch": round(t_multi * 1e6, 1),
                "us_per_step": round(t_multi / best.steps_per_launch * 1e6, 1),
                "effective_GBs": round(8 * padded * best.steps_per_launch / t_multi / 1e9, 1),
                "note": "effective = 8 B per padded node and step over the launch time; each "
                        "launch reads and writes every field once for all its steps "
                        "(LDS-resident tiles), so this exceeds the one-step HBM roofline"}),
            "copy_of_packed_buffer": {"bytes": 8 * int(a.numel()),
                                      "us": round(t_copy * 1e6, 1),
                                      "GBs": round(8 * a.numel() / t_copy / 1e9, 1)},
            "one_step_kernel": {"tile_config": {"rows_per_lane": one_cfg.rows_per_lane,
                                                "waves": one_cfg.waves, "blocks": one_cfg.blocks,
                                                "xcd_slots": one_cfg.xcd_slots},
                                "us": round(t_kernel * 1e6, 1),
                                "GBs": round(8 * padded / t_kernel / 1e9, 1),
                                "ms_per_step_with_boundary": round(ens.timings[one], 4)},
            "note": "batched: the autotuned form (several steps per launch on LDS-resident "
                    "tiles with the boundary inside, or boundary + step kernels per step), "
                    "graph-replayed; per-instance: HeatStep per grid, eager"}


def heat_config_timing(dev, steps=1000, cpu_steps=100):
    """Heat workload (256^2, 8 seeded Gaussian bumps, Dirichlet 0 walls,
    dt = 0.2 dx^2 / alpha, 1000 steps): the fused HeatStep loop (one launch
    per step, the boundary folded into the step) replayed as a HIP graph of 10
    steps, the reference-shaped module loop (legacy GridBoundary + Laplacian +
    ForwardEuler, eager), the C FP32 restatement single-threaded on
    ``cpu_steps`` steps, and the accumulated error of the GPU state against the
    FP64 oracle at steps 10, 100 and 1000 (relative to max |u|)."""
    import torch
    from oracle import c_oracle
    from oracle import np_oracle as O
    from pde_module_amd import ForwardEuler, HeatStep, Laplacian
    from pde_module_amd.deprecated import GridBoundary as LegacyGridBoundary
    from pde_module_amd.workloads import heat_2d_config
    cfg = heat_2d_config(seed=0)
    g = cfg.g
    u0 = torch.from_numpy(cfg.u0).to(dev)
    mk_bc = lambda: LegacyGridBoundary(u0, cfg.dx, g)
    bc = mk_bc()
    bc.dirichlet_BC("ALL", 0.0)
    fused = HeatStep(bc, cfg.dx, g, cfg.alpha, cfg.dt)
    u = fused(fused(u0))                       # both ping-pong buffers set up
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    a = fused._bufs[0]
    per = 10                                   # steps per replay (even: back in buffer a)
    with torch.cuda.graph(graph):
        x = a
        for _ in range(per):
            x = fused(x)
    torch.cuda.synchronize()
    a.copy_(u0)
    snaps = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps // per):
        graph.replay()
        if per * (k + 1) in (10, 100, 1000):
            torch.cuda.synchronize()
            snaps[per * (k + 1)] = a.cpu().numpy()
    torch.cuda.synchronize()
    t_fused = (time.perf_counter() - t0) / steps
    # several steps per launch on LDS-resident tiles (HeatStep.run): the whole
    # 1000-step loop captured as one graph (84 launches of up to 12 steps)
    multi = HeatStep(mk_bc(), cfg.dx, g, cfg.alpha, cfg.dt)
    multi.boundary.dirichlet_BC("ALL", 0.0)
    ms = multi.run(u0.clone(), steps, steps_per_launch=MULTI_SPL)   # plan, buffers, warm-up
    torch.cuda.synchronize()
    ma = multi._bufs[0]
    mgraph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(mgraph):
        mout = multi.run(ma, steps, steps_per_launch=MULTI_SPL)
    ma.copy_(u0)
    torch.cuda.synchronize()
    mgraph.replay()
    torch.cuda.synchronize()
    same_as_fused = bool(torch.equal(mout.view(torch.int32), a.view(torch.int32)))
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        ma.copy_(u0)
        mgraph.replay()
    torch.cuda.synchronize()
    t_multi = (time.perf_counter() - t0) / (reps * steps)
    mplan = multi._mplan[0][4], multi._mplan[1][1]
    # module-by-module (the reference's loop shape), eager, 100 steps
    bcm = mk_bc()
    bcm.dirichlet_BC("ALL", 0.0)
    lap, euler = Laplacian(1, cfg.dx, g), ForwardEuler()
    um = u0
    for _ in range(3):
        u2 = bcm(um)
        um = euler(u2, lap(u2, cfg.alpha), cfg.dt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(100):
        u2 = bcm(um)
        um = euler(u2, lap(u2, cfg.alpha), cfg.dt)
    torch.cuda.synchronize()
    t_mod = (time.perf_counter() - t0) / 100
    # CPU: the C restatement, one thread
    shape = cfg.u0.shape[:3]
    bc_fn = lambda x: c_oracle.bc_legacy(x, bc.boundary_ijk_indices, bc.boundary_interior,
                                         bc.boundary_type, bc.boundary_value, cfg.dx)
    w = O.laplacian_weights(cfg.dx)
    uc = cfg.u0.copy()
    t0 = time.perf_counter()
    for _ in range(cpu_steps):
        u2c = bc_fn(uc)
        uc = c_oracle.forward_euler(u2c, c_oracle.laplacian(u2c, w, np.float32(cfg.alpha), g),
                                    np.float32(cfg.dt))
    t_cpu = (time.perf_counter() - t0) / cpu_steps
    # FP64 oracle: accumulated error
    ijk, _, inter = O.legacy_boundary_layout(shape, g)
    bt, bv = np.ones((len(ijk), 1), np.int8), np.zeros((len(ijk), 1))
    bc64 = lambda x: O.bc_legacy(x, ijk, inter, bt, bv, cfg.dx)
    w64 = O.laplacian_weights(cfg.dx, np.float64)
    mask = O.interior_mask(shape, g)[..., None]
    u64 = cfg.u0.astype(np.float64)
    errs = {}
    for step in range(1, steps + 1):
        u64 = O.heat_step(u64, w64, cfg.alpha, g, cfg.dt, bc64)
        if step in snaps:
            s = snaps[step]
            errs[str(step)] = float(np.max(np.abs(s[mask] - u64[mask])) / np.max(np.abs(u64[mask])))
    lups = cfg.n * cfg.n
    return {"grid": [cfg.n, cfg.n], "steps": steps,
            "multistep_graph_us_per_step": round(t_multi * 1e6, 2),
            "multistep_glups": round(lups / t_multi / 1e9, 2),
            "multistep": {"steps_per_launch": MULTI_SPL, "region": list(mplan[0]),
                          "tiles": int(mplan[1]), "state_after_1000_steps_bitwise_equal_"
                          "to_one_step_per_launch": same_as_fused},
            "fused_graph_us_per_step": round(t_fused * 1e6, 2),
            "fused_glups": round(lups / t_fused / 1e9, 2),
            "module_loop_us_per_step": round(t_mod * 1e6, 2),
            "cpu_1thread_us_per_step": round(t_cpu * 1e6, 1),
            "cpu_glups": round(lups / t_cpu / 1e9, 4),
            "rel_error_vs_fp64": errs,
            "note": "cache-resident (266 KB field): launch- and latency-bound, not an HBM "
                    "figure; multistep: several steps per launch on LDS-resident tiles "
                    "(pde_heat_multistep_f32), fused: one launch per step"}


def projection_cavity_timing(dev, n=1024, steps=2000, batch_steps=20, seeds=16,
                             checkpoint_every=10):
    """Projection-method cavity workload (1024^2, Re = 100, 50 Jacobi sweeps per
    step, seeded divergence-free u0 of amplitude 1e-3).  One instance over all
    ``steps`` steps: the graph-replayed forward, then the kinetic-energy
    gradient dJ/du0 through the transposed steps (CavityRollout: every step's
    velocity kept, 17 GB at 2000 steps, and checkpoints every
    ``checkpoint_every`` steps with recompute), each timed after a warm-up run
    that captures the graphs.  Batch: ``seeds`` instances (seeds 0..15) as 16
    instances in one captured graph, and as one ProjectionCavityBatch (forward
    and batched gradient, every state kept), timed per step over
    ``batch_steps`` steps (every state of 16 seeds for 2000 steps would be
    270 GB)."""
    import torch
    from pde_module_amd.flow import CavityRollout, ProjectionCavity, ProjectionCavityBatch
    from pde_module_amd.workloads import ns_cavity_u0
    cav = ProjectionCavity(n, jacobi_sweeps=50, device=dev)
    u0 = torch.from_numpy(ns_cavity_u0(0, n)).to(dev)
    p0 = torch.zeros(cav.shape + (1,), device=dev)
    area = float(cav.dx * cav.dx)
    out = {"grid": [n, n], "jacobi_sweeps": 50, "dt": cav.dt, "steps": steps}
    for name, K in (("store_all", None), ("checkpointed", checkpoint_every)):
        ro = CavityRollout(cav, steps, checkpoint_every=K, use_graph=True)
        for rep in range(2):                     # first run captures the graphs
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            uN, _ = ro.forward_states(u0, p0)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            gu, _ = ro.backward_states(area * uN * cav.mask, torch.zeros_like(p0))
            torch.cuda.synchronize()
            t2 = time.perf_counter()
        fwd, bwd = (t1 - t0) / steps, (t2 - t1) / steps
        if K is None:
            out["fwd_ms_per_step"] = round(fwd * 1e3, 3)
        out[name] = {"checkpoint_every": K, "fwd_plus_dKE_du0_ms_per_step": round((fwd + bwd) * 1e3, 3),
                     "ratio_to_fwd": round((fwd + bwd) / out["fwd_ms_per_step"] / 1e-3, 2)}
        del ro
    # 16 seeds as 16 separate instances in one graph (one cavity's work buffers)
    us = [torch.from_numpy(ns_cavity_u0(sd, n)).to(dev) for sd in range(seeds)]
    ps = [torch.zeros(cav.shape + (1,), device=dev) for _ in range(seeds)]
    for u, p in zip(us, ps):
        cav.step(u, p, u, p)
    torch.cuda.synchronize()
    gr = torch.cuda.CUDAGraph()
    with torch.cuda.graph(gr):
        for u, p in zip(us, ps):
            cav.step(u, p, u, p)
    gr.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(batch_steps):
        gr.replay()
    torch.cuda.synchronize()
    out["16_instances_one_graph_ms_per_step"] = round((time.perf_counter() - t0) / batch_steps
                                                      * 1e3, 3)
    del gr, us, ps
    # the same 16 seeds as one ProjectionCavityBatch (every instance's sweeps in
    # one launch per sweep), forward and the batched gradient
    bat = ProjectionCavityBatch(n, seeds, jacobi_sweeps=50, device=dev)
    ub = torch.from_numpy(np.stack([ns_cavity_u0(sd, n) for sd in range(seeds)])).to(dev)
    pb = torch.zeros((seeds,) + cav.shape + (1,), device=dev)
    ro = CavityRollout(bat, batch_steps, checkpoint_every=None, use_graph=True)
    for rep in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        uN, _ = ro.forward_states(ub, pb)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ro.backward_states(area * uN * cav.mask, torch.zeros_like(pb))
        torch.cuda.synchronize()
        t2 = time.perf_counter()
    fwd, bwd = (t1 - t0) / batch_steps, (t2 - t1) / batch_steps
    out["batch_steps"] = batch_steps
    out["batch_of_16_ms_per_step"] = round(fwd * 1e3, 3)
    out["batch_of_16_fwd_plus_dKE_du0_ms_per_step"] = round((fwd + bwd) * 1e3, 3)
    out["batch_speedup_vs_16_instances"] = round(out["16_instances_one_graph_ms_per_step"] /
                                                 out["batch_of_16_ms_per_step"], 2)
    out["note"] = ("graph-replayed forward and transposed steps; gradient = dJ/du0 of the "
                   "interior kinetic energy after `steps` steps; batch = 16 seeds "
                   "(ProjectionCavityBatch, store-all rollout), measured")
    return out


def ldc_module_timing(dev, n=201, steps=20000, eager_steps=2000, beta=0.3, nu=0.01, cfl=0.5):
    """The reference's own benchmark harness shape (Benchmark/Functional_BC/
    LDC_Graph_Capture.py and LDC_No_Graph.py): the artificial-compressibility
    lid-driven cavity, n^2 nodes, Re = 100, built from the drop-in modules
    (legacy GridBoundary, Laplacian, Grad, Divergence, the mat-vec glue,
    ForwardEuler) exactly as the script composes them; one step captured as
    a HIP graph and replayed `steps` times, and `eager_steps` steps launched
    eagerly.  The reference publishes no numbers for it."""
    import torch
    from pde_module_amd import ForwardEuler
    from pde_module_amd.deprecated import GridBoundary as LegacyGridBoundary
    from pde_module_amd.FDM import Divergence, Grad, Laplacian, functional
    from pde_module_amd.mesh import UniformGridMesh, create_structured_field
    from pde_module_amd.stencil import ElementWise
    dx = 1.0 / (n - 1)
    g = 1
    grid = UniformGridMesh(dx, nodes_per_axis=(n, n, 1), origin=(0.0, 0.0, 0.0), ghost_cells=g)
    dt = min(cfl * float(dx ** 2 / nu), cfl * dx / (1 + 1 / beta))
    u = create_structured_field(grid, "node", 2, device=dev)
    p = create_structured_field(grid, "node", 1, device=dev)
    u_BC = LegacyGridBoundary(u, dx, g)
    u_BC.dirichlet_BC("ALL", 0.0)
    u_BC.dirichlet_BC("+Y", 1.0, 0)
    p_BC = LegacyGridBoundary(p, dx, g)
    p_BC.vonNeumann_BC("ALL", 0.0)
    u_lapl, u_grad = Laplacian(2, dx, g), Grad(2, u.shape, dx, ghost_cells=g)
    u_div, p_grad = Divergence(2, u.shape, dx, g), Grad(1, p.shape, dx, g)
    u_step, p_step = ForwardEuler(), ForwardEuler()
    conv = ElementWise("matvec")
    u_F = torch.empty_like(u)

    def step(u, p):
        u_fix = u_BC(u)
        u_diff = u_lapl(u_fix, nu)
        du = u_grad(u_fix)
        u_incomp = u_div(u_fix, -beta)
        u_conv = conv(du, u)
        p_fix = p_BC(p)
        dp = p_grad(p_fix, -1.0)
        functional.linear_combination(u_F, (1, dp), (-1, u_conv), (1, u_diff))
        return u_step(u_fix, u_F, dt), p_step(p_fix, u_incomp, dt)

    u, p = step(u, p)                      # set-up step
    for _ in range(20):
        u, p = step(u, p)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(eager_steps):
        u, p = step(u, p)
    torch.cuda.synchronize()
    t_eager = (time.perf_counter() - t0) / eager_steps
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        step(u, p)                         # the state lives in the Euler outputs
    for _ in range(20):
        graph.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        graph.replay()
    torch.cuda.synchronize()
    t_graph = (time.perf_counter() - t0) / steps
    uh = u[1:-1, 1:-1].cpu()
    return {"grid": [n, n], "steps_graph": steps, "steps_eager": eager_steps,
            "graph_ms_per_step": round(t_graph * 1e3, 4),
            "eager_ms_per_step": round(t_eager * 1e3, 4),
            "graph_total_s": round(t_graph * steps, 3),
            "finite": bool(torch.isfinite(uh).all()),
            "note": "the reference's Functional_BC benchmark shape (201^2, 20 000 captured "
                    "steps) on the drop-in modules, one launch per module call"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    import pde_module_amd as pm
    from pde_module_amd import _native as NV
    from pde_module_amd.FDM import functional as F

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.only_wave:
        print(json.dumps({"wave2048_fwd_vs_fwd_adjoint":
                          wave_adjoint_timing(dev, steps=args.wave_steps)}))
        return
    if args.only_heat:
        print(json.dumps({"heat_ensemble_512": heat_ensemble_timing(dev),
                          "heat_256": heat_config_timing(dev)}))
        return
    if args.only_ldc:
        print(json.dumps({"ldc_201_modules": ldc_module_timing(dev)}))
        return
    if args.only_cavity:
        print(json.dumps({"projection_cavity_1024":
                          projection_cavity_timing(dev, steps=args.cavity_steps)}))
        return
    gen = torch.Generator(device=dev).manual_seed(0)
    s = torch.rand((P, P, P, 1), device=dev, generator=gen) * 2 - 1
    v = torch.rand((P, P, P, 3), device=dev, generator=gen) * 2 - 1
    periodic = pm.PeriodicBoundary(G)
    periodic(s)
    periodic(v)
    lap = pm.Laplacian(1, 1.0 / N, G)
    plap = pm.PeriodicLaplacian(1, 1.0 / N, G)
    grad = pm.Grad(1, (P, P, P), 1.0 / N, G)
    div = pm.Divergence(3, (P, P, P), 1.0 / N, G)

    def step():
        return plap(s, 1.0)

    def dropin_step():
        periodic(s)
        return lap(s, 1.0)

    # correctness of the timed step on this grid: the one-launch periodic
    # Laplacian equals ghost fill + the drop-in Laplacian on every interior point
    want = dropin_step().clone()
    got = step()
    inner = (slice(G, -G),) * 3
    if not torch.equal(got[inner].view(torch.int32), want[inner].view(torch.int32)):
        raise AssertionError("periodic Laplacian differs from fill + Laplacian")
    del want
    # the GPU's clocks ramp up over its first ~10 ms of work after idle: run the
    # step for PREWARM_S seconds (untimed) before the W warm-up steps, so the
    # K timed steps measure the steady state a long solver loop runs at
    t_end = time.perf_counter() + PREWARM_S
    while time.perf_counter() < t_end:
        for _ in range(10):
            step()
        torch.cuda.synchronize()
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed / args.steps * 1e3
    glups = LUPS / (ms_per_step * 1e-3) / 1e9

    def event_time(fn, reps):
        """Average duration of fn's launches, HIP events on the launch stream."""
        stream = torch.cuda.current_stream(dev)
        for _ in range(3):
            fn()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(reps):
            fn()
        b.record(stream)
        b.synchronize()
        return a.elapsed_time(b) / reps * 1e-3

    reps = max(10, args.steps)
    t_step_kernel = event_time(step, reps)
    step_gbs = LAP_BYTES * LUPS / t_step_kernel / 1e9
    t_lap = event_time(lambda: lap(s, 1.0), reps)
    lap_gbs = LAP_BYTES * LUPS / t_lap / 1e9
    t_dropin = event_time(dropin_step, reps)
    out_l = torch.empty_like(s)
    out_g = torch.empty_like(v)
    out_d = torch.empty_like(s)
    t_grad = event_time(lambda: grad(s, 1.0), reps // 2)
    t_div = event_time(lambda: div(v, 1.0), reps // 2)
    t_fused = event_time(lambda: F.fused_lap_grad_div(s, v, out_l, out_g, out_d, 1.0 / N, G),
                         reps // 2)
    t_copy = event_time(lambda: NV.call("pde_copy_f32", s.data_ptr(), out_l.data_ptr(),
                                        s.numel(), NV.stream_of(s)), reps // 2)
    copy_gbs = 2 * 4 * s.numel() / t_copy / 1e9
    # ForwardEuler over every padded scalar (read cur and k, write out: 12 B)
    euler = pm.ForwardEuler()
    t_euler = event_time(lambda: euler(s, out_l, 1e-3), reps // 2)
    # functional GridBoundary, Dirichlet 0 on every face: the module call (copy
    # of the field + the boundary writes) and the same writes in place
    gbc = pm.GridBoundary(s, 1.0 / N, G)
    gbc.dirichlet_BC("ALL", 0.0)
    t_bc = event_time(lambda: gbc(s), reps // 2)
    bc_field = s.clone()
    t_bc_inplace = event_time(lambda: gbc.apply_inplace(bc_field), reps // 2)
    del bc_field

    progress(f"operators timed: step {t_step_kernel * 1e6:.1f} us, drop-in laplacian "
             f"{t_lap * 1e6:.1f} us")
    wave = ensemble = cavity = heat = ldc = None
    if not args.quick:
        wave = wave_adjoint_timing(dev, steps=args.wave_steps)
        progress("wave forward / forward+adjoint timed")
        wave["receivers_reached"] = wave_adjoint_timing(dev, steps=args.wave_steps,
                                                        rec_distance=400, variants=False)
        progress("wave gradient with receivers the front reaches: checked against recompute")
        if not args.no_cpu:
            wave["cpu_forward_step"] = wave_cpu_timing(args.cpu_seconds / 2)
            progress("wave forward step timed on the host (C restatement)")
        ensemble = heat_ensemble_timing(dev)
        progress("heat ensemble timed")
        cavity = projection_cavity_timing(dev)
        progress("projection cavity timed")
        heat = heat_config_timing(dev)
        progress("heat 256^2 timed")
        ldc = ldc_module_timing(dev)
        progress("module-path cavity (the reference's benchmark harness) timed")
    traffic = pmc_traffic()
    rec = {
        "metric": "GLUPS & % HBM peak (3D 7-pt FP32); fwd+adjoint vs fwd step time",
        "value": round(glups, 2),
        "unit": "GLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: s ~ U[-1,1] (torch.Generator seed 0), periodic ghosts",
        "prewarm_s": PREWARM_S,
        "config": {"workload": "3D 7-point Laplacian, 512^3 periodic grid (514^3 padded, g=1), "
                               "step = one periodic Laplacian launch (wrapped neighbour "
                               "indices; equal bit for bit to periodic ghost fill + Laplacian)",
                   "grid_padded": [P, P, P], "interior_points": LUPS,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(step_gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(step_gbs / HBM_PEAK_GBS, 4),
                     "traffic": traffic.get("lap3d_ring<true, 2, 3, 4>") if traffic else None,
                     "kernel": "lap3d_ring<true, 2, 3, 4> (periodic Laplacian)",
                     "avg_launch_us": round(t_step_kernel * 1e6, 2),
                     "algorithmic_bytes_per_launch": LAP_BYTES * LUPS},
        "extra": {
            "dropin_laplacian_kernel": {
                "kernel": "lap3d_ring<false, 2, 3, 4>", "us": round(t_lap * 1e6, 2),
                "GBs": round(lap_gbs, 1), "frac": round(lap_gbs / HBM_PEAK_GBS, 4),
                "glups": round(LUPS / t_lap / 1e9, 1),
                "traffic": traffic.get("lap3d_ring<false, 2, 3, 4>") if traffic else None},
            "dropin_step_fill_plus_laplacian": {
                "us": round(t_dropin * 1e6, 2), "glups": round(LUPS / t_dropin / 1e9, 1)},
            "grad": {"us": round(t_grad * 1e6, 1), "GBs": round(GRAD_BYTES * LUPS / t_grad / 1e9, 1),
                     "traffic": traffic.get("ops3d_band<2, false, 8, 2>") if traffic else None},
            "div": {"us": round(t_div * 1e6, 1), "GBs": round(DIV_BYTES * LUPS / t_div / 1e9, 1),
                    "kernel": "div3d_ring<2, 8>",
                    "traffic": traffic.get("div3d_ring<2, 8>") if traffic else None},
            "fused_lap_grad_div": {"us": round(t_fused * 1e6, 1),
                                   "GBs": round(FUSED_BYTES * LUPS / t_fused / 1e9, 1),
                                   "kernels": ["ops3d_band<3, false, 8, 2>", "div3d_ring<2, 8>"],
                                   "traffic": (traffic["ops3d_band<3, false, 8, 2>"]
                                               + traffic["div3d_ring<2, 8>"]
                                               if traffic and "div3d_ring<2, 8>" in traffic
                                               and "ops3d_band<3, false, 8, 2>" in traffic
                                               else None),
                                   "vs_separate": round((t_lap + t_grad + t_div) / t_fused, 2)},
            "copy_GBs": round(copy_gbs, 1),
            "forward_euler": {"us": round(t_euler * 1e6, 1),
                              "GBs": round(12 * s.numel() / t_euler / 1e9, 1),
                              "frac": round(12 * s.numel() / t_euler / 1e9 / HBM_PEAK_GBS, 4),
                              "note": "out = cur + dt*k over the 514^3 padded array, 12 B per "
                                      "scalar"},
            "grid_boundary_dirichlet_all": {
                "module_call_us": round(t_bc * 1e6, 1),
                "module_call_GBs": round(8 * s.numel() / t_bc / 1e9, 1),
                "in_place_us": round(t_bc_inplace * 1e6, 1),
                "ops": int(gbc.table.nops),
                "note": "functional GridBoundary, Dirichlet 0 on every face of 514^3: the "
                        "module call copies the field (8 B per scalar) and applies the writes; "
                        "in place only the O(surface) writes"},
            "wave2048_fwd_vs_fwd_adjoint": wave,
            "heat_ensemble_512": ensemble,
            "projection_cavity_1024": cavity,
            "heat_256": heat,
            "ldc_201_modules": ldc,
        },
        "fwd_adjoint_vs_fwd": wave["ratio"] if wave else None,   # checkpointed, reversed
    }
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_seconds)
        threads = max(cpu)
        rec["cpu_baseline"] = {"value": round(cpu[1], 4), "unit": "GLUPS", "cores": 1,
                               "kind": "port",
                               "sample": "C FP32 restatement of the reference Laplacian "
                                         "(oracle/pde_oracle.c, -O2 -ffp-contract=off), "
                                         "32-plane slabs of the 514^3 grid, single thread",
                               "multicore": {"value": round(cpu[threads], 4), "cores": threads},
                               "host_cpu": host_cpu_model()}
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
