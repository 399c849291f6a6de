"""ORACLE — test infrastructure only (tests/, smoke() and bench.py's cpu_baseline).

NumPy restatement of PDE_Module's uniform-grid FDM hot path, parametrised on the
floating type:

* ``dtype=np.float32`` — a second, independent FP32 restatement.  It keeps the
  reference's operation order (NumPy rounds every + and * separately), so it must
  agree bit for bit with the C restatement in ``pde_oracle.c``; the two are
  written differently (vectorised slices here, scalar loops there), so an
  indexing mistake in either shows up as a mismatch.
* ``dtype=np.float64`` — the FP64 oracle used to report accumulated error.

Fields are arrays of padded shape ``(X, Y, Z, C)`` (vector) or ``(X, Y, Z, N, D)``
(matrix), C-order, components innermost (reference
``pm/mesh/mesh_functions.py:108-159``).  Operator outputs have NaN ghosts here, so
any path that reads an uninitialised ghost shows up as NaN in the interior.

Pinned by tests/test_oracle_*.py: closed forms (quadratics, discrete
eigenfunctions, whole-loop heat and standing-wave solutions), scipy brute force
(sparse Kronecker sums, ``ndimage.correlate``), ``numpy.gradient``, summation by
parts, and the boundary identities (``(ghost + adj) / 2 == v`` etc.).
"""
from __future__ import annotations

import numpy as np


# --------------------------------------------------------------------------- geometry
def eligible_axes(shape3):
    """Axes with size > 1, ascending (reference ``pm/FDM/utils.py:24-27``)."""
    return tuple(a for a, s in enumerate(shape3) if s > 1)


def interior(shape3, g, axis=None, shift=0):
    """Slices of the interior nodes, optionally shifted by ``shift`` along ``axis``
    (reference ``pm/FDM/kernel/laplacian.py:54-57``: node = tid + g on eligible
    axes)."""
    out = []
    for a, p in enumerate(shape3):
        lo, hi = (g, p - g) if p > 1 else (0, p)
        if a == axis:
            lo, hi = lo + shift, hi + shift
        out.append(slice(lo, hi))
    return tuple(out)


def interior_mask(shape3, g):
    m = np.zeros(shape3, dtype=bool)
    m[interior(shape3, g)] = True
    return m


# --------------------------------------------------------------------------- weights
def laplacian_weights(dx, dtype=np.float32):
    """Reference ``pm/FDM/module/laplacian.py:32-35``: FP64 values cast to the
    field's float type."""
    return tuple(dtype(v) for v in (1.0 / dx**2, -2.0 / dx**2, 1.0 / dx**2))


def central_weights(dx, dtype=np.float32):
    """Reference ``pm/FDM/module/grad.py:37-39`` and ``divergence.py:36-39``."""
    return tuple(dtype(v) for v in (-1.0 / (2 * dx), 0.0, 1.0 / (2 * dx)))


# --------------------------------------------------------------------------- operators
def laplacian(u, w, alpha, g):
    """Reference ``pm/FDM/kernel/laplacian.py:34-59``; u is (X, Y, Z, C)."""
    dt = u.dtype.type
    shape3 = u.shape[:3]
    out = np.full_like(u, np.nan)
    acc = np.zeros_like(u[interior(shape3, g)])
    for a in eligible_axes(shape3):
        for s, ws in zip((-1, 0, 1), w):
            acc = acc + dt(ws) * u[interior(shape3, g, a, s)]
    out[interior(shape3, g)] = dt(alpha) * acc
    return out


def grad(u, w, alpha, g, force_matrix=False):
    """Reference ``pm/FDM/kernel/grad.py:78-100``.  u is (X, Y, Z, N); returns
    (X, Y, Z, D) when N == 1 and not force_matrix, else (X, Y, Z, N, D)."""
    dt = u.dtype.type
    shape3 = u.shape[:3]
    axes = eligible_axes(shape3)
    N, D = u.shape[3], len(axes)
    out = np.full(shape3 + (N, D), np.nan, dtype=u.dtype)
    inner = interior(shape3, g)
    for t, a in enumerate(axes):
        acc = np.zeros_like(u[inner])
        for s, ws in zip((-1, 0, 1), w):
            acc = acc + dt(ws) * u[interior(shape3, g, a, s)]
        out[inner + (slice(None), t)] = dt(alpha) * acc
    if N == 1 and not force_matrix:
        return out.reshape(shape3 + (D,))
    return out


def divergence(u, w, alpha, g, matrix_in=False):
    """Reference ``pm/FDM/kernel/divergence.py:85-106``.  Vector input
    (X, Y, Z, D) -> (X, Y, Z, 1); matrix input (X, Y, Z, N, D) -> (X, Y, Z, N)."""
    dt = u.dtype.type
    shape3 = u.shape[:3]
    axes = eligible_axes(shape3)
    m = u if matrix_in else u.reshape(shape3 + (1, u.shape[3]))
    rows = m.shape[3]
    out = np.full(shape3 + (rows,), np.nan, dtype=u.dtype)
    inner = interior(shape3, g)
    acc = np.zeros(m[inner].shape[:-1], dtype=u.dtype)
    for t, a in enumerate(axes):
        for s, ws in zip((-1, 0, 1), w):
            acc = acc + dt(ws) * m[interior(shape3, g, a, s) + (slice(None), t)]
    out[inner] = dt(alpha) * acc
    return out


def forward_euler(cur, k, dt_step):
    """Reference ``pm/time_step/forwardEuler.py:79``: cur + dt * k everywhere."""
    return cur + cur.dtype.type(dt_step) * k


def rk1_step(f, bc, t, dt, *fields):
    """Forward Euler over several fields as the reference's ``RungeKatta1``
    documents it (``pm/time_step/rungeKatta.py:21-25,63-83``): bc, then
    y + dt * f(t, y).  (The reference's code steps by dt / 2, ``:80``, and
    never runs: its setup raises, ``:59``.)"""
    ys = bc(t, *fields)
    ks = f(t, *ys)
    return tuple(forward_euler(y, k, dt) for y, k in zip(ys, ks))


def rk2_step(f, bc, t, dt, *fields):
    """Midpoint step of the reference's ``RungeKatta2.forward``
    (``pm/time_step/rungeKatta.py:145-174``): y0 = bc(t, y) (kept as a copy),
    k1 = f(t, y0), y1 = bc(t, y0 + dt/2 k1), k2 = f(t + dt/2, y1),
    y0 + dt k2.  dt / 2 and t + dt / 2 are Python (FP64) arithmetic, rounded
    to FP32 where the Euler update uses them."""
    y0 = tuple(np.array(y, copy=True) for y in bc(t, *fields))
    k1 = f(t, *y0)
    y1 = tuple(forward_euler(y, k, dt / 2) for y, k in zip(y0, k1))
    y1 = bc(t, *y1)
    k2 = f(t + dt / 2, *y1)
    return tuple(forward_euler(y, k, dt) for y, k in zip(y0, k2))


def damping_layer(u, grid_shape, num_layers, p, points, scale, farfield=None, out=None):
    """Sponge / viscous damping layer, node by node as the reference kernels do
    (``pm/FDM/kernel/farfield.py:36-77``, ``viscousDampingLayer.py:36-74``):
    axis = argmin over axes of min(i, s - 1 - i) (size-1 axes excluded, first
    axis on ties); dx from the near face's layer limit; factor =
    scale * (dx / num_layers) ** p with the quotient in the field's precision
    and the power in FP64 rounded once; out = -factor * (u - farfield), or
    -factor * u without a far field.  Other nodes of ``out`` keep their values
    (zeros for a fresh output, as the modules zero theirs at setup)."""
    dt = u.dtype.type
    out = np.zeros_like(u) if out is None else out
    gs = [int(s) for s in grid_shape]
    limits = [(num_layers, (s - 1) - num_layers) if s > 1 else (0, 0) for s in gs]
    for pt in np.asarray(points).reshape(-1, 3):
        dists = [min(int(pt[i]), gs[i] - int(pt[i]) - 1) if gs[i] > 1 else np.iinfo(np.int32).max
                 for i in range(3)]
        axis = int(np.argmin(dists))
        x = int(pt[axis])
        if (gs[axis] - 1 - x) < x:
            dx = x - limits[axis][1]
        else:
            dx = limits[axis][0] - x
        base = dt(dx) / dt(num_layers)
        prof = dt(np.float64(base) ** np.float64(np.float32(p)))
        f = dt(scale) * prof
        idx = tuple(int(i) for i in pt)
        v = u[idx] if farfield is None else u[idx] - np.asarray(farfield, dtype=u.dtype).reshape(u[idx].shape)
        out[idx] = (-f) * v
    return out


def immersed_boundary(u, bitmask, btype, bvalue, dx, fill_value=0.0):
    """Staircase immersed boundary, node by node (reference
    ``pm/FDM/boundary/immersedBoundary.py:177-300`` with the host steps
    ``:54-152``): solid nodes = nonzero bitmask in C order; a solid node is a
    boundary node when a neighbour +-1 along an eligible axis is fluid (0);
    btype / bvalue are the (boundary node, component) tables in that order.
    Dirichlet: out = v.  Neumann: running mean over the fluid neighbours
    (axes ascending, side -1 then +1) of u[nb] - (s * v) * dx.  Interior solid
    nodes take fill_value.  Everything else is u.  Neighbours are flat index
    +- the axis stride (the reference indexes without a bounds check, so on the
    outermost layer a row index -1 is the previous row's last node); a
    neighbour outside the whole array raises ValueError."""
    dt = u.dtype.type
    out = np.array(u, copy=True)
    shape3 = bitmask.shape
    elig = [a for a in range(3) if shape3[a] > 1]
    strides = (shape3[1] * shape3[2], shape3[2], 1)
    bflat = bitmask.reshape(-1)
    C = u.shape[3]
    uflat = u.reshape(-1, C)
    oflat = out.reshape(-1, C)
    # neighbours by flat index +- axis stride, the reference's unchecked indexing;
    # an index outside the whole array (which NumPy would wrap) is an error, as
    # in the product's ImmersedBoundary.finalize
    solid = [int(f) for f in np.flatnonzero(bflat)]
    for f in solid:
        for a in elig:
            for s in (-1, 1):
                if not 0 <= f + s * strides[a] < bflat.size:
                    raise ValueError("a solid node's neighbour falls outside the array")
    bnd, inner = [], []
    for f in solid:
        is_b = any(bflat[f + s * strides[a]] == 0 for a in elig for s in (-1, 1))
        (bnd if is_b else inner).append(f)
    for b, f in enumerate(bnd):
        for c in range(C):
            v = dt(bvalue[b, c])
            if btype[b, c] == 1:
                oflat[f, c] = v
            elif btype[b, c] == 2:
                avg, n = dt(0.0), dt(1.0)
                for a in elig:
                    for s in (-1, 1):
                        q = f + s * strides[a]
                        if bflat[q] == 0:
                            cval = uflat[q, c] - (dt(s) * v) * dt(dx)
                            avg = avg + (cval - avg) / n
                            n = n + dt(1.0)
                oflat[f, c] = avg
    for f in inner:
        oflat[f] = dt(fill_value)
    return out


def jacobi(p, rhs, g, h2, inv):
    """One Jacobi sweep on the interior of a scalar field: acc over eligible
    axes ascending of p[n - e] + p[n + e] (accumulated from 0, left to right),
    then inv * (acc - h2 * rhs) (restates pde_jacobi_f32; the pressure step of
    the projection cavity, no reference counterpart)."""
    dt = p.dtype.type
    shape3 = p.shape[:3]
    inner = interior(shape3, g)
    acc = np.zeros_like(p[inner])
    for a in eligible_axes(shape3):
        acc = acc + p[interior(shape3, g, a, -1)]
        acc = acc + p[interior(shape3, g, a, +1)]
    out = np.full_like(p, np.nan)
    out[inner] = dt(inv) * (acc - dt(h2) * rhs[inner])
    return out


def projection_step(u, p, lw, cw, nu, dt_step, h2, inv, sweeps, g, ubc, pbc):
    """One step of the projection cavity (the composition
    pde_module_amd.flow.ProjectionCavity.step runs): ubc / pbc map a field to
    its boundary-applied copy."""
    dt = u.dtype.type
    u2 = ubc(u)
    du = grad(u2, cw, dt(1.0), g)
    conv = matvec(du, u2)
    diff = laplacian(u2, lw, dt(nu), g)
    F = linear_combination((-1, conv), (1, diff))
    us = forward_euler(u2, F, dt_step)
    us2 = ubc(us)
    rhs = divergence(us2, cw, dt(1.0 / dt_step), g)
    for _ in range(sweeps):
        p = jacobi(pbc(p), rhs, g, h2, inv)
    gp = grad(pbc(p), cw, dt(1.0), g)
    return forward_euler(us2, gp, -dt_step), p


# --------------------------------------------------------------------------- boundary groups
def no_ghost_shape(shape3, g):
    """Reference ``ExplicitUniformGridStencil.py:134-148``."""
    return tuple(p - 2 * g if p > 1 else p for p in shape3)


def legacy_boundary_layout(shape3, g):
    """Boundary node list, named groups and interior directions of the legacy
    ``GridBoundary`` (reference ``pm/deprecated/gridBoundary.py:60-116``).

    Returns (ijk (nb, 3) int32 sorted lexicographically, groups dict name -> ids,
    interior (nb, 3) int32)."""
    ng = no_ghost_shape(shape3, g)
    parts = []
    for i in range(3):
        if ng[i] == 1:
            continue
        for lim in (0, ng[i] - 1):
            side = list(ng)
            side[i] = 1
            idx = np.moveaxis(np.indices(side, dtype=np.int32), 0, -1).reshape(-1, 3)
            idx[:, i] = lim
            for ax in range(3):
                if ng[ax] != 1:
                    idx[:, ax] += g
            parts.append(idx)
    ijk = np.unique(np.concatenate(parts, axis=0), axis=0).astype(np.int32)
    ids = np.arange(len(ijk), dtype=np.int32)
    groups = {"ALL": ids}
    for axis, name in enumerate("XYZ"):
        if ng[axis] == 1:
            continue
        for lim, parity in ((g, "-"), (ng[axis] - 1 + g, "+")):
            groups[parity + name] = ids[ijk[:, axis] == lim]
    inter = np.zeros((len(ijk), 3), dtype=np.int32)
    for axis, name in enumerate("XYZ"):
        if ng[axis] == 1:
            continue
        for parity, sign in (("-", 1), ("+", -1)):
            inter[groups[parity + name], axis] = sign
    return ijk, groups, inter


def functional_groups(shape3, g):
    """Named face groups of the functional ``GridBoundary`` as flat C-order node
    ids (reference ``pm/FDM/boundary/functional_boundary.py:111-137``)."""
    flat = np.arange(int(np.prod(shape3)), dtype=np.int32).reshape(shape3)
    groups, parts = {}, []
    for i, name in enumerate("XYZ"):
        if shape3[i] <= 1:
            continue
        for prefix, sign in (("-", -1), ("+", 1)):
            sl = [slice(None)] * 3
            sl[i] = g if sign == -1 else -(g + 1)
            for x, p in enumerate(shape3):
                if p > 1 and x != i:
                    sl[x] = slice(g, -g) if g > 0 else slice(None)
            groups[prefix + name] = flat[tuple(sl)].ravel()
            parts.append(groups[prefix + name])
    groups["ALL"] = np.unique(np.concatenate(parts)).astype(np.int32)
    return groups


# --------------------------------------------------------------------------- boundary apply
def _record_value(grp, t, dt):
    """A group's record value at time t in the field's precision: constant,
    ramp min(t / tmax, 1) * const (reference ``boundary_functions.py:85-89``)
    or sine amp * sin(omega * t) (the examples' ``sin_wave``,
    ``examples/wave2d/Interference.py:53-62``; the sine of the rounded product
    evaluated in FP64 and rounded once)."""
    if grp.get("sine") is not None:
        amp, omega = grp["sine"]
        arg = dt(omega) * dt(t)
        return dt(amp) * dt(np.sin(np.float64(arg)))
    v = dt(grp["const"])
    if grp.get("tmax") is not None:
        ramp = dt(t) / dt(grp["tmax"])
        ramp = min(ramp, dt(1.0))
        v = dt(ramp) * dt(grp["const"])
    return v


def bc_legacy(u, ijk, inter, btype, bvalue, dx, func_groups=(), t=0.0, coords=None,
              params=None):
    """Legacy table apply (reference ``pm/deprecated/gridBoundary.py:267-302``).
    btype/bvalue are (nb, C).  Writes of one launch never collide, so applying
    them vectorised per axis is the same as the reference's threads.

    Then each function group in order, one thread per entry (reference
    ``pm/deprecated/func_boundary.py:96-134``, launched after the table at
    ``gridBoundary.py:244-245``): dicts with ids (entry ids), type, var, name
    and either record keys (const / tmax / sine) or ``func`` called per entry
    as func(u, nodeID (3,), var, coords, t, dx, params[name]) -> scalar, with
    every read from the unmodified u."""
    dt = u.dtype.type
    X, Y, Z, C = u.shape
    sn = np.array([Y * Z, Z, 1], dtype=np.int64)
    out = u.copy()
    fo, fi = out.reshape(-1, C), u.reshape(-1, C)
    node = ijk.astype(np.int64) @ sn
    for a in range(3):
        b = np.nonzero(inter[:, a] != 0)[0]
        s = inter[b, a].astype(np.int64)
        ghost, adj = node[b] - s * sn[a], node[b] + s * sn[a]
        k = np.where(s > 0, dt(-2.0), dt(2.0)).astype(u.dtype) * dt(dx)
        for c in range(C):
            ty, v = btype[b, c], bvalue[b, c].astype(u.dtype)
            d = ty == 1
            fo[node[b][d], c] = v[d]
            fo[ghost[d], c] = dt(2.0) * v[d] - fi[adj[d], c]
            n = ty == 2
            fo[ghost[n], c] = k[n] * v[n] + fi[adj[n], c]
    params = params or {}
    for grp in func_groups:
        var = int(grp["var"])
        for e in np.asarray(grp["ids"], dtype=np.int64):
            nid = ijk[e].astype(np.int64)
            if grp.get("func") is not None:
                val = dt(grp["func"](u, nid, var, coords, dt(t), dt(dx),
                                     params.get(grp.get("name"), 0)))
            else:
                val = _record_value(grp, t, dt)
            for a in range(3):
                s = int(inter[e, a])
                if s == 0:
                    continue
                f = int(nid @ sn)
                ghost, adj = f - s * sn[a], f + s * sn[a]
                if grp["type"] == 1:
                    fo[f, var] = val
                    fo[ghost, var] = dt(2.0) * val - fi[adj, var]
                else:
                    kk = (dt(-2.0) if s > 0 else dt(2.0)) * dt(dx)
                    fo[ghost, var] = kk * val + fi[adj, var]
    return out


def bc_functional(u, g, groups, dx, t=0.0, coords=None, params=None):
    """Functional apply (reference ``pm/FDM/boundary/functional_boundary.py:260-400``).

    ``groups`` is a list of dicts with keys ids (flat node ids), type (1 Dirichlet,
    2 Neumann), comps (selected components), and the value: const, tmax (None
    for constant) or sine (amp, omega), the node vector being the node's own
    with the selected components replaced (``boundary_functions.py:42-45``);
    or ``func``, called per node as func(u, nodeID (3,), coords, t, dx,
    params[name]) -> vector (C,), with every read from the unmodified u
    (``functional_boundary.py:376``)."""
    dt = u.dtype.type
    X, Y, Z, C = u.shape
    shape3 = (X, Y, Z)
    sn = np.array([Y * Z, Z, 1], dtype=np.int64)
    out = u.copy()
    fo, fi = out.reshape(-1, C), u.reshape(-1, C)
    params = params or {}
    for grp in groups:
        ids = np.asarray(grp["ids"], dtype=np.int64)
        comps = list(grp["comps"])
        if grp.get("func") is not None:
            idx3 = np.stack([ids // (Y * Z), (ids // Z) % Y, ids % Z], axis=1)
            val = np.array([np.broadcast_to(np.asarray(
                grp["func"](u, nid, coords, dt(t), dt(dx), params.get(grp.get("name"), 0)),
                dtype=u.dtype), (C,)) for nid in idx3], dtype=u.dtype).reshape(len(ids), C)
        else:
            val = fi[ids].copy()
            val[:, comps] = _record_value(grp, t, dt)
        if grp["type"] == 1:
            fo[ids] = val
        idx = np.stack([ids // (Y * Z), (ids // Z) % Y, ids % Z], axis=1)
        for a in eligible_axes(shape3):
            s = np.where(idx[:, a] == g, 1, np.where(idx[:, a] == shape3[a] - 1 - g, -1, 0))
            sel = s != 0
            q, sq = ids[sel], s[sel]
            ghost, adj = q - sq * sn[a], q + sq * sn[a]
            k = np.where(sq > 0, dt(-2.0), dt(2.0)).astype(u.dtype) * dt(dx)
            for c in comps:
                vc = val[sel, c]
                if grp["type"] == 1:
                    fo[ghost, c] = dt(2.0) * vc - fi[adj, c]
                else:
                    fo[ghost, c] = k * vc + fi[adj, c]
    return out


def periodic_fill(u, g):
    """In-place periodic ghost fill, axis by axis (no counterpart in the
    reference)."""
    shape3 = u.shape[:3]
    for a in eligible_axes(shape3):
        E = shape3[a] - 2 * g
        for i in range(g):
            dst = [slice(None)] * 3
            src = [slice(None)] * 3
            dst[a], src[a] = i, i + E
            u[tuple(dst)] = u[tuple(src)]
            dst[a], src[a] = shape3[a] - g + i, g + i
            u[tuple(dst)] = u[tuple(src)]
    return u


# --------------------------------------------------------------------------- composed steps
def heat_step(u, w, alpha, g, dt_step, bc):
    """One step of the reference heat loop (``examples/2D_Heat_Equation.py:58-62``):
    u2 = BC(u); L = lap(u2, alpha); u_next = u2 + dt * L.  ``bc`` maps u -> u2.
    Ghosts of u_next are NaN (the reference leaves them uninitialised)."""
    u2 = bc(u)
    lap = laplacian(u2, w, alpha, g)
    return forward_euler(u2, lap, dt_step)


def wave_step(u, v, w, c2, g, dt_step, bc, first):
    """One step of the reference u/v wave loop (``examples/wave2d/standingWave.py:72-76``)
    with the aliasing it relies on: from the second step on v is updated in place
    before u reads it (symplectic Euler); the first step reads the old v.
    ``c2`` is a scalar or a per-node (X, Y, Z, 1) array multiplying the Laplacian."""
    dt = u.dtype.type
    u2 = bc(u)
    lap = laplacian(u2, w, dt(1.0) if np.ndim(c2) else c2, g)
    if np.ndim(c2):
        lap = c2 * lap
    v_next = forward_euler(v, lap, dt_step)
    u_next = forward_euler(u2, v if first else v_next, dt_step)
    return u_next, v_next


# --------------------------------------------------------------------------- node-wise glue
# Element-wise operations over every padded node (ghosts included), as the
# reference's element-wise launches cover the whole array.
def outer(a, b):
    """wp.outer per node: (..., la), (..., lb) -> (..., la, lb), M[i][j] = a[i]*b[j]
    (reference ``pm/FDM/module/elementWiseOps.py:18``)."""
    return a[..., :, None] * b[..., None, :]


def scalar_vector_multiply(s, b):
    """s[0] * b per node, s a vec(1) field (reference
    ``pm/FDM/kernel/elementWiseOps.py:17-21``)."""
    sc = s.reshape(s.shape[:3] + (1,) * (b.ndim - 3))
    return sc * b


def divide_by_scalar(m, rho):
    """m / rho[0] per node (reference ``examples/LDC/LDC_RK2.py:117``)."""
    return m / rho.reshape(rho.shape[:3] + (1,) * (m.ndim - 3))


def matvec(m, v):
    """Warp mat * vec per node: out[i] = m[i][0]*v[0] + m[i][1]*v[1] + ...,
    columns ascending (the array product ``du * u``,
    ``examples/LDC/LDC_No_Graph.py:87``)."""
    acc = m[..., :, 0] * v[..., None, 0]
    for j in range(1, m.shape[-1]):
        acc = acc + m[..., :, j] * v[..., None, j]
    return acc


def linear_combination(*terms):
    """c0*x0 + c1*x1 + ... left to right (the array-operator chain
    ``dp - u_conv + u_diff``, ``examples/LDC/LDC_No_Graph.py:90``)."""
    (c0, x0), rest = terms[0], terms[1:]
    dt = x0.dtype.type
    acc = dt(c0) * x0
    for c, x in rest:
        acc = acc + dt(c) * x
    return acc


def shifted_scale(x, scale, shift, comp=-1):
    """scale * (x - shift) on component comp (all when comp < 0)
    (``get_p_op``, ``examples/LDC/LDC_RK2.py:111-116``)."""
    dt = x.dtype.type
    y = dt(scale) * (x - dt(shift))
    if comp < 0:
        return y
    out = x.copy()
    out[..., comp] = y[..., comp]
    return out


# --------------------------------------------------------------------------- adjoints
# Transposes of the operators above, seen as linear maps from the whole padded
# field to its interior nodes (output ghosts are not part of the range).  Each
# is written as a gather in a fixed order: for every padded node m, over axes
# ascending and s = -1, 0, +1, the term w_s * g[m - s e_a] of every interior
# node m - s e_a that reads m.  Pinned by dot-product tests against the
# forward functions (tests/test_oracle_adjoints.py).
def _shift(a, axis, s):
    """out[m] = a[m - s e_axis], zero where m - s e_axis leaves the array."""
    out = np.zeros_like(a)
    n = a.shape[axis]
    src = [slice(None)] * a.ndim
    dst = [slice(None)] * a.ndim
    if s > 0:
        dst[axis], src[axis] = slice(s, n), slice(0, n - s)
    elif s < 0:
        dst[axis], src[axis] = slice(0, n + s), slice(-s, n)
    out[tuple(dst)] = a[tuple(src)]
    return out


def _interior_only(g, g_cells):
    """g with everything outside the interior nodes set to 0 (the transposes
    never read output ghosts)."""
    out = np.zeros_like(g)
    inner = interior(g.shape[:3], g_cells)
    out[inner] = g[inner]
    return out


def laplacian_adjoint(gout, w, alpha, g):
    dt = gout.dtype.type
    gz = _interior_only(gout, g)
    acc = np.zeros_like(gz)
    for a in eligible_axes(gz.shape[:3]):
        for s, ws in zip((-1, 0, 1), w):
            acc = acc + dt(ws) * _shift(gz, a, s)
    return dt(alpha) * acc


def grad_adjoint(gout, w, alpha, g):
    """gout (X,Y,Z,N,D) or (X,Y,Z,D) for a scalar input; returns (X,Y,Z,N)."""
    dt = gout.dtype.type
    shape3 = gout.shape[:3]
    axes = eligible_axes(shape3)
    m = gout if gout.ndim == 5 else gout.reshape(shape3 + (1, len(axes)))
    gz = _interior_only(m, g)
    acc = np.zeros(shape3 + (m.shape[3],), dtype=gout.dtype)
    for t, a in enumerate(axes):
        for s, ws in zip((-1, 0, 1), w):
            acc = acc + dt(ws) * _shift(gz[..., t], a, s)
    return dt(alpha) * acc


def divergence_adjoint(gout, w, alpha, g, D):
    """gout (X,Y,Z,rows); returns (X,Y,Z,rows,D) (rows == 1: reshape to a vec D
    field for the vector form)."""
    dt = gout.dtype.type
    shape3 = gout.shape[:3]
    axes = eligible_axes(shape3)
    gz = _interior_only(gout, g)
    out = np.zeros(shape3 + (gout.shape[3], D), dtype=gout.dtype)
    for t, a in enumerate(axes):
        acc = np.zeros_like(gz)
        for s, ws in zip((-1, 0, 1), w):
            acc = acc + dt(ws) * _shift(gz, a, s)
        out[..., t] = dt(alpha) * acc
    return out


def boundary_linear_adjoint(gout, writes):
    """Transpose of the linear part of a boundary apply.  ``writes`` lists the
    final (dst, src, sign) of every written scalar (sign 0 for a constant
    write): out[dst] = sign * in[src] + const, out = in elsewhere."""
    fo = gout.reshape(-1)
    gin = fo.copy()
    dsts = np.array([d for d, _, _ in writes], dtype=np.int64)
    gin[dsts] = 0
    for d, s, sign in writes:
        if sign != 0:
            gin[s] = gin[s] + gout.dtype.type(sign) * fo[d]
    return gin.reshape(gout.shape)


def heat_step_adjoint(gnext, w, alpha, g, dt_step):
    """Transpose of u2 -> u_next = u2 + dt * (alpha * lap(u2)) (interior) /
    u2 (ghosts): g_u2 = g + dt * (alpha * accT), accT the Laplacian transpose
    sum of g over interior nodes."""
    dt = gnext.dtype.type
    gz = _interior_only(gnext, g)
    acc = np.zeros_like(gz)
    for a in eligible_axes(gz.shape[:3]):
        for s, ws in zip((-1, 0, 1), w):
            acc = acc + dt(ws) * _shift(gz, a, s)
    return gnext + dt(dt_step) * (dt(alpha) * acc)


def _acc_full(u, w, g):
    """Laplacian sum (no alpha) on interior nodes, 0 elsewhere."""
    dt = u.dtype.type
    shape3 = u.shape[:3]
    out = np.zeros_like(u)
    acc = np.zeros_like(u[interior(shape3, g)])
    for a in eligible_axes(shape3):
        for s, ws in zip((-1, 0, 1), w):
            acc = acc + dt(ws) * u[interior(shape3, g, a, s)]
    out[interior(shape3, g)] = acc
    return out


def wave_step_core(u2, v, c2, w, alpha, g, dt_step, first):
    """The fused u/v wave update on a boundary-applied u2 (no boundary call):
    interior vn = v + dt*(c2*(alpha*acc)), un = u2 + dt*(v if first else vn);
    ghosts vn = v, un = u2.  c2 None means 1."""
    dt = u2.dtype.type
    M = interior_mask(u2.shape[:3], g)[..., None]
    rhs = dt(alpha) * _acc_full(u2, w, g)
    if c2 is not None:
        rhs = c2 * rhs
    vn = np.where(M, v + dt(dt_step) * rhs, v)
    un = np.where(M, u2 + dt(dt_step) * (v if first else vn), u2)
    return un, vn


def wave_step_adjoint(gun, gvn, c2, u2, w, alpha, g, dt_step, first):
    """Transpose of wave_step_core in (u2, v) and its derivative in c2.
    Returns (gu2, gv, dc2) with dc2 the increment added to the c2 gradient."""
    dt = gun.dtype.type
    M = interior_mask(gun.shape[:3], g)[..., None]
    d = dt(dt_step)
    gvt = np.where(M, gvn if first else gvn + d * gun, gvn)
    h = np.where(M, gvt if c2 is None else c2 * gvt, dt(0))
    acc = np.zeros_like(h)
    for a in eligible_axes(h.shape[:3]):
        for s, ws in zip((-1, 0, 1), w):
            acc = acc + dt(ws) * _shift(h, a, s)
    gu2 = gun + d * (dt(alpha) * acc)
    gv = np.where(M, (gvn + d * gun) if first else gvt, gvn)
    dc2 = np.where(M, d * (gvt * (dt(alpha) * _acc_full(u2, w, g))), dt(0))
    return gu2, gv, dc2
