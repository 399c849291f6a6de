"""Writes the committed golden fixtures (tests/golden/*.npz).

The reference ships no golden vectors for this path and cannot run here (its
kernels need warp-lang, which is not installed), so these fixtures come from
this build's oracle: the C FP32 restatement for the expected FP32 outputs and
the FP64 NumPy oracle for the FP64 references.  The oracle itself is pinned by
tests/test_oracle_pins.py; the fixtures freeze its outputs so any later change
to the oracle or to the GPU path shows up as a diff.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import c_oracle as C  # noqa: E402
from oracle import np_oracle as O  # noqa: E402


def rnd(rng, shape):
    return rng.uniform(-1, 1, shape).astype(np.float32)


def operators(rng):
    out = {}
    cases = [("lap2d_v1", (10, 12, 1), 1), ("lap3d_v1", (6, 7, 8), 1), ("lap3d_v3", (6, 7, 8), 3)]
    for name, shape, c in cases:
        u = rnd(rng, shape + (c,))
        dx, alpha = 0.1, 0.7
        out[name] = dict(u=u, dx=dx, alpha=alpha, g=1,
                         f32=C.laplacian(u, O.laplacian_weights(dx), np.float32(alpha), 1),
                         f64=O.laplacian(u.astype(np.float64), O.laplacian_weights(dx, np.float64), alpha, 1))
    u = rnd(rng, (10, 12, 1, 1))
    out["grad2d_v1"] = dict(u=u, dx=0.05, alpha=1.5, g=1,
                            f32=C.grad(u, O.central_weights(0.05), np.float32(1.5), 1),
                            f64=O.grad(u.astype(np.float64), O.central_weights(0.05, np.float64), 1.5, 1))
    u = rnd(rng, (10, 12, 1, 2))
    out["grad2d_v2"] = dict(u=u, dx=0.05, alpha=1.5, g=1,
                            f32=C.grad(u, O.central_weights(0.05), np.float32(1.5), 1),
                            f64=O.grad(u.astype(np.float64), O.central_weights(0.05, np.float64), 1.5, 1))
    v = rnd(rng, (10, 12, 1, 2))
    out["div2d_v2"] = dict(u=v, dx=0.05, alpha=-0.5, g=1,
                           f32=C.divergence(v, O.central_weights(0.05), np.float32(-0.5), 1),
                           f64=O.divergence(v.astype(np.float64), O.central_weights(0.05, np.float64), -0.5, 1))
    m = rnd(rng, (10, 12, 1, 2, 2))
    out["div2d_m22"] = dict(u=m, dx=0.05, alpha=2.0, g=1,
                            f32=C.divergence(m, O.central_weights(0.05), np.float32(2.0), 1, True),
                            f64=O.divergence(m.astype(np.float64), O.central_weights(0.05, np.float64), 2.0, 1, True))
    return out


def boundaries(rng):
    out = {}
    shape = (9, 8, 1)
    fg = O.functional_groups(shape, 1)
    func_cases = {
        "func_dirichlet_all": (1, [dict(ids=fg["ALL"], type=1, comps=[0], const=0.0)]),
        "func_lid": (2, [dict(ids=fg["ALL"], type=1, comps=[0, 1], const=0.0),
                         dict(ids=fg["+Y"], type=1, comps=[0], const=1.0)]),
        "func_neumann_all": (1, [dict(ids=fg["ALL"], type=2, comps=[0], const=0.5)]),
        "func_mixed": (3, [dict(ids=fg["ALL"], type=1, comps=[0, 2], const=1.25),
                           dict(ids=fg["-X"], type=2, comps=[1], const=-2.0),
                           dict(ids=fg["+Y"], type=1, comps=[2], const=3.0, tmax=0.5)]),
    }
    for name, (c, groups) in func_cases.items():
        u = rnd(rng, shape + (c,))
        rec = dict(u=u, dx=0.1, t=0.2, f32=C.bc_functional(u, 1, groups, 0.1, t=0.2))
        for i, gdef in enumerate(groups):
            rec[f"g{i}_ids"] = np.asarray(gdef["ids"], np.int32)
            rec[f"g{i}_meta"] = np.array([gdef["type"], gdef["const"], gdef.get("tmax") or 0.0])
            rec[f"g{i}_comps"] = np.asarray(gdef["comps"], np.int32)
        out[name] = rec
    ijk, groups, inter = O.legacy_boundary_layout(shape, 1)
    for name, c, setter in [
        ("legacy_default", 1, lambda bt, bv: None),
        ("legacy_lid", 2, lambda bt, bv: (bv.__setitem__((groups["+Y"], 0), 1.0))),
        ("legacy_neumann_all", 1, lambda bt, bv: (bt.__setitem__(slice(None), 2), bv.__setitem__(slice(None), 0.5))),
        ("legacy_mixed", 3, lambda bt, bv: (bt.__setitem__((groups["-X"], 1), 2),
                                           bv.__setitem__((groups["-X"], 1), -2.0),
                                           bv.__setitem__((slice(None), 0), 1.25))),
    ]:
        u = rnd(rng, shape + (c,))
        bt = np.ones((len(ijk), c), np.int8)
        bv = np.zeros((len(ijk), c), np.float32)
        setter(bt, bv)
        out[name] = dict(u=u, dx=0.1, ijk=ijk, inter=inter, btype=bt, bvalue=bv,
                         f32=C.bc_legacy(u, ijk, inter, bt, bv, 0.1))
    return out


def loops(rng):
    out = {}
    n, g, alpha = 32, 1, 0.01
    dx = 1.0 / (n - 1)
    dt = 0.2 * dx * dx / alpha
    xs = np.arange(-1, n + 1) * dx
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    u0 = (np.exp(-((X - 0.4) ** 2 + (Y - 0.6) ** 2) / 0.02)).reshape(n + 2, n + 2, 1, 1)
    ijk, _, inter = O.legacy_boundary_layout(u0.shape[:3], g)
    bt, bv = np.ones((len(ijk), 1), np.int8), np.zeros((len(ijk), 1), np.float32)
    u32, u64 = u0.astype(np.float32), u0.astype(np.float64)
    w32, w64 = O.laplacian_weights(dx), O.laplacian_weights(dx, np.float64)
    for _ in range(50):
        u2 = C.bc_legacy(u32, ijk, inter, bt, bv, dx)
        u32 = C.forward_euler(u2, C.laplacian(u2, w32, np.float32(alpha), g), np.float32(dt))
        u64 = O.heat_step(u64, w64, alpha, g, dt, lambda x: O.bc_legacy(x, ijk, inter, bt, bv, dx))
    out["heat32_50"] = dict(u0=u0.astype(np.float32), dx=dx, alpha=alpha, dt=dt, steps=50,
                            f32=u32, f64=u64)
    n = 64
    dx = 1.0 / (n - 1)
    dt = 0.5 * dx
    xs = np.arange(-1, n + 1) * dx
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    u0 = (np.sin(2 * np.pi * X) * np.sin(np.pi * Y)).reshape(n + 2, n + 2, 1, 1).astype(np.float32)
    ijk, _, inter = O.legacy_boundary_layout(u0.shape[:3], g)
    bt, bv = np.ones((len(ijk), 1), np.int8), np.zeros((len(ijk), 1), np.float32)
    w32 = O.laplacian_weights(dx)
    u, v = u0, np.zeros_like(u0)
    for i in range(100):
        u2 = C.bc_legacy(u, ijk, inter, bt, bv, dx)
        nv = C.forward_euler(v, C.laplacian(u2, w32, np.float32(1.0), g), np.float32(dt))
        u = C.forward_euler(u2, v if i == 0 else nv, np.float32(dt))
        v = nv
    out["wave64_100"] = dict(u0=u0, dx=dx, c2=1.0, dt=dt, steps=100, f32_u=u, f32_v=v)
    s = rnd(rng, (18, 18, 18, 1))
    vv = rnd(rng, (18, 18, 18, 3))
    C.periodic_fill(s, 1)
    C.periodic_fill(vv, 1)
    dxp = 1.0 / 16
    out["fused16"] = dict(s=s, v=vv, dx=dxp,
                          lap=C.laplacian(s, O.laplacian_weights(dxp), np.float32(1.0), 1),
                          grad=C.grad(s, O.central_weights(dxp), np.float32(1.0), 1),
                          div=C.divergence(vv, O.central_weights(dxp), np.float32(1.0), 1))
    return out


def main():
    rng = np.random.default_rng(20261203)
    for group, fn in (("operators", operators), ("boundaries", boundaries), ("loops", loops)):
        cases = fn(rng)
        flat = {}
        for case, rec in cases.items():
            for k, val in rec.items():
                flat[f"{case}/{k}"] = np.asarray(val)
        path = os.path.join(HERE, f"{group}.npz")
        np.savez_compressed(path, **flat)
        print(f"wrote {path} ({os.path.getsize(path)} bytes, {len(cases)} cases)")


if __name__ == "__main__":
    main()
