"""ORACLE — test infrastructure only (tests/ and bench.py's cpu_baseline).

FP64 torch restatement of the heat and u/v wave loops with the legacy
GridBoundary's default (Dirichlet 0 on every boundary entry, reference
``pm/deprecated/gridBoundary.py:267-302``, ``pm/FDM/boundary/boundary.py:59-65``),
written with differentiable torch ops so that ``torch.autograd`` gives the
reference gradient the GPU adjoints are checked against.  The reference itself
has no autodiff.  Pinned against the NumPy FP64 oracle in
tests/test_oracle_torch.py (same forward values to round-off) and by central
finite differences of its own forward.
"""
from __future__ import annotations

import torch


def _inner(P, g):
    return slice(g, P - g) if P > 1 else slice(None)


def bc_dirichlet0(u, g):
    """Legacy default boundary on every face: boundary nodes 0 and face
    ghosts -u[adj]; every read from the input u."""
    X, Y, Z = u.shape[:3]
    P = (X, Y, Z)
    axes = [a for a in range(3) if P[a] > 1]
    out = u.clone()
    for a in axes:
        for ghost, adj in ((g - 1, g + 1), (P[a] - g, P[a] - 2 - g)):
            dst = [_inner(P[b], g) for b in range(3)]
            src = list(dst)
            dst[a], src[a] = ghost, adj
            out[tuple(dst)] = -u[tuple(src)]
    for a in axes:
        for node in (g, P[a] - 1 - g):
            dst = [_inner(P[b], g) for b in range(3)]
            dst[a] = node
            out[tuple(dst)] = 0.0
    return out


def lap_acc(u, dx, g):
    """Laplacian sum (no alpha) on interior nodes, 0 elsewhere."""
    P = tuple(u.shape[:3])
    w = (1.0 / dx**2, -2.0 / dx**2, 1.0 / dx**2)
    inner = tuple(_inner(p, g) for p in P)
    acc = torch.zeros_like(u[inner])
    for a in range(3):
        if P[a] <= 1:
            continue
        for s, ws in zip((-1, 0, 1), w):
            sl = list(inner)
            sl[a] = slice(g + s, P[a] - g + s)
            acc = acc + ws * u[tuple(sl)]
    out = torch.zeros_like(u)
    out[inner] = acc
    return out


def interior_mask(u, g):
    m = torch.zeros(u.shape[:3], dtype=torch.bool)
    m[tuple(_inner(p, g) for p in u.shape[:3])] = True
    return m[..., None]


def heat_rollout(u0, dx, g, alpha, dt, steps):
    u = u0
    M = interior_mask(u0, g)
    for _ in range(steps):
        u2 = bc_dirichlet0(u, g)
        u = torch.where(M, u2 + dt * (alpha * lap_acc(u2, dx, g)), u2)
    return u


def wave_rollout(c2, u0, v0, dx, g, dt, steps, alpha=1.0, src_idx=None, src_wavelet=None,
                 rec_idx=None):
    """Returns (u_N, v_N, records) with the same per-step order as the GPU
    WaveRollout: boundary, source into v, fused u/v update, sample u."""
    u, v = u0, v0
    M = interior_mask(u0, g)
    recs = []
    for n in range(steps):
        u2 = bc_dirichlet0(u, g)
        if src_idx is not None:
            v = v.reshape(-1).index_add(0, src_idx, dt * src_wavelet[n]).reshape(v.shape)
        rhs = alpha * lap_acc(u2, dx, g)
        if c2 is not None:
            rhs = c2 * rhs
        vn = torch.where(M, v + dt * rhs, v)
        un = torch.where(M, u2 + dt * (v if n == 0 else vn), u2)
        u, v = un, vn
        if rec_idx is not None:
            recs.append(u.reshape(-1)[rec_idx])
    rec = torch.stack(recs) if recs else None
    return u, v, rec
