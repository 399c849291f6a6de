"""Committed golden fixtures (tests/golden/*.npz, written by make_golden.py).

CPU: the NumPy FP32 restatement (independent of the C code that wrote the
fixtures) reproduces every FP32 fixture bit for bit, and the FP64 oracle its
FP64 fixtures.  GPU: the HIP modules reproduce the FP32 fixtures bit for bit.
"""
import os

import numpy as np
import pytest
import torch

from conftest import same_bits
from helpers import ghost_free_mask
from oracle import np_oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(group):
    z = np.load(os.path.join(HERE, f"{group}.npz"), allow_pickle=False)
    cases = {}
    for key in z.files:
        case, field = key.split("/", 1)
        cases.setdefault(case, {})[field] = z[key]
    return cases


def func_groups(rec):
    groups, i = [], 0
    while f"g{i}_ids" in rec:
        ty, const, tmax = rec[f"g{i}_meta"]
        groups.append(dict(ids=rec[f"g{i}_ids"], type=int(ty), comps=list(rec[f"g{i}_comps"]),
                           const=float(const), tmax=float(tmax) if tmax > 0 else None))
        i += 1
    return groups


def np_operator(name, rec, dtype):
    u = rec["u"].astype(dtype)
    dx, alpha, g = float(rec["dx"]), float(rec["alpha"]), int(rec["g"])
    if name.startswith("lap"):
        return O.laplacian(u, O.laplacian_weights(dx, dtype), alpha, g)
    if name.startswith("grad"):
        return O.grad(u, O.central_weights(dx, dtype), alpha, g)
    return O.divergence(u, O.central_weights(dx, dtype), alpha, g, matrix_in=name.endswith("m22"))


@pytest.mark.parametrize("name", sorted(load("operators")))
def test_operator_fixtures_on_cpu(name):
    rec = load("operators")[name]
    assert same_bits(np_operator(name, rec, np.float32), rec["f32"])
    np.testing.assert_array_equal(np_operator(name, rec, np.float64), rec["f64"])


@pytest.mark.parametrize("name", sorted(load("boundaries")))
def test_boundary_fixtures_on_cpu(name):
    rec = load("boundaries")[name]
    if name.startswith("func"):
        got = O.bc_functional(rec["u"], 1, func_groups(rec), np.float32(rec["dx"]), t=float(rec["t"]))
    else:
        got = O.bc_legacy(rec["u"], rec["ijk"], rec["inter"], rec["btype"], rec["bvalue"],
                          np.float32(rec["dx"]))
    assert same_bits(got, rec["f32"])


def test_loop_fixtures_on_cpu():
    cases = load("loops")
    h = cases["heat32_50"]
    u = h["u0"]
    shape3 = u.shape[:3]
    ijk, _, inter = O.legacy_boundary_layout(shape3, 1)
    bt, bv = np.ones((len(ijk), 1), np.int8), np.zeros((len(ijk), 1), np.float32)
    dx, alpha, dt = float(h["dx"]), float(h["alpha"]), float(h["dt"])
    for _ in range(int(h["steps"])):
        u = O.heat_step(u, O.laplacian_weights(dx), alpha, 1, dt,
                        lambda x: O.bc_legacy(x, ijk, inter, bt, bv, np.float32(dx)))
    mask = ghost_free_mask(shape3, 1, (1,))
    assert same_bits(u, h["f32"], mask)
    err = np.max(np.abs(h["f32"][mask] - h["f64"][mask])) / np.max(np.abs(h["f64"][mask]))
    assert err < 1e-6
    w = cases["wave64_100"]
    u, v = w["u0"], np.zeros_like(w["u0"])
    shape3 = u.shape[:3]
    ijk, _, inter = O.legacy_boundary_layout(shape3, 1)
    bt, bv = np.ones((len(ijk), 1), np.int8), np.zeros((len(ijk), 1), np.float32)
    dx, dt = float(w["dx"]), float(w["dt"])
    for i in range(int(w["steps"])):
        u, v = O.wave_step(u, v, O.laplacian_weights(dx), np.float32(1.0), 1, dt,
                           lambda x: O.bc_legacy(x, ijk, inter, bt, bv, np.float32(dx)), first=(i == 0))
    mask = ghost_free_mask(shape3, 1, (1,))
    assert same_bits(u, w["f32_u"], mask) and same_bits(v, w["f32_v"], mask)


# --------------------------------------------------------------------- GPU
@pytest.mark.gpu
def test_fixtures_on_gpu(dev):
    import pde_module_amd as pm
    from pde_module_amd.deprecated import GridBoundary as LegacyGridBoundary
    from pde_module_amd.FDM import functional as F

    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    for name, rec in load("operators").items():
        print("fixture", name, flush=True)
        u = T(rec["u"])
        dx, alpha, g = float(rec["dx"]), float(rec["alpha"]), int(rec["g"])
        shape3 = tuple(u.shape[:3])
        if name.startswith("lap"):
            out = pm.Laplacian(int(u.shape[3]), dx, g)(u, alpha)
        elif name.startswith("grad"):
            out = pm.Grad(int(u.shape[3]), shape3, dx, g)(u, alpha)
        elif name.endswith("m22"):
            out = pm.Divergence(tuple(u.shape[3:]), shape3, dx, g)(u, alpha)
        else:
            out = pm.Divergence(int(u.shape[3]), shape3, dx, g)(u, alpha)
        want = rec["f32"]
        assert same_bits(out.cpu().numpy().reshape(want.shape), want,
                         ghost_free_mask(shape3, g, want.shape[3:])), name
    for name, rec in load("boundaries").items():
        print("fixture", name, flush=True)
        u = T(rec["u"])
        if name.startswith("func"):
            bc = pm.GridBoundary(u, float(rec["dx"]), 1)
            from pde_module_amd.FDM.boundary import get_ramp_func
            for grp in func_groups(rec):
                val = get_ramp_func(grp["const"], grp["tmax"]) if grp["tmax"] else grp["const"]
                bc.set_BC(np.asarray(grp["ids"]), val, grp["type"], [int(c) for c in grp["comps"]])
            got = bc(u, float(rec["t"]))
        else:
            bc = LegacyGridBoundary(u, float(rec["dx"]), 1)
            bc.boundary_type[...] = rec["btype"]
            bc.boundary_value[...] = rec["bvalue"]
            got = bc(u)
        assert same_bits(got.cpu().numpy(), rec["f32"]), name
    cases = load("loops")
    h = cases["heat32_50"]
    u = T(h["u0"])
    bc = LegacyGridBoundary(u, float(h["dx"]), 1)
    step = pm.HeatStep(bc, float(h["dx"]), 1, float(h["alpha"]), float(h["dt"]))
    for _ in range(int(h["steps"])):
        u = step(u)
    mask = ghost_free_mask(tuple(u.shape[:3]), 1, (1,))
    assert same_bits(u.cpu().numpy(), h["f32"], mask)
    w = cases["wave64_100"]
    u, v = T(w["u0"]), torch.zeros_like(T(w["u0"]))
    bc = LegacyGridBoundary(u, float(w["dx"]), 1)
    step = pm.WaveStep(bc, float(w["dx"]), 1, 1.0, float(w["dt"]))
    for _ in range(int(w["steps"])):
        u, v = step(u, v)
    mask = ghost_free_mask(tuple(u.shape[:3]), 1, (1,))
    assert same_bits(u.cpu().numpy(), w["f32_u"], mask)
    assert same_bits(v.cpu().numpy(), w["f32_v"], mask)
    f = cases["fused16"]
    s, vv = T(f["s"]), T(f["v"])
    lap, grad, div = torch.zeros_like(s), torch.zeros_like(vv), torch.zeros_like(s)
    F.fused_lap_grad_div(s, vv, lap, grad, div, float(f["dx"]), 1)
    mask3 = ghost_free_mask((18, 18, 18), 1, ())
    for got, want in ((lap, f["lap"]), (grad, f["grad"]), (div, f["div"])):
        assert same_bits(got.cpu().numpy()[mask3], want[mask3])
