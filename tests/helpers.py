"""Test-side helpers (not part of the package)."""
import numpy as np

from pde_module_amd._native import BC_COPY, BC_GHOST_D, BC_SET, BCV_ARRAY, BCV_SINE, bc_slots


def apply_ops_host(u, ops, t=0.0, values=None):
    """Interpret a resolved boundary op table on the host in FP32, to check the
    table resolution against the oracle without a GPU.  ``values``: the time
    array's function values (t[1:]) for array-valued ops."""
    fin = np.ascontiguousarray(u, dtype=np.float32).ravel()
    out = fin.copy()
    t = np.float32(t)
    for op in ops:
        v = op["value"]
        if op["vkind"] == BCV_SINE:
            v = op["value"] * np.float32(np.sin(np.float64(op["tmax"] * t)))
        elif op["vkind"] == BCV_ARRAY:
            v = np.float32(values[int(bc_slots(np.array([op]))[0])])
        elif op["tmax"] > 0:
            v = min(t / op["tmax"], np.float32(1.0)) * op["value"]
        k = op["kind"]
        if k == BC_SET:
            out[op["dst"]] = v
        elif k == BC_COPY:
            out[op["dst"]] = fin[op["src"]]
        elif k == BC_GHOST_D:
            out[op["dst"]] = np.float32(2.0) * v - fin[op["src"]]
        else:
            out[op["dst"]] = op["coef"] * v + fin[op["src"]]
    return out.reshape(u.shape)


def ghost_free_mask(shape3, g, comps_shape=()):
    """True on non-ghost nodes (broadcast over element axes)."""
    m = np.zeros(shape3, dtype=bool)
    sl = tuple(slice(g, p - g) if p > 1 else slice(None) for p in shape3)
    m[sl] = True
    return np.broadcast_to(m.reshape(shape3 + (1,) * len(comps_shape)), shape3 + tuple(comps_shape))


def face_ghost_mask(shape3, g):
    """Non-ghost nodes plus the face ghosts of the first ghost layer (the ghosts
    a 3-point stencil reads); edge and corner ghosts excluded."""
    m = ghost_free_mask(shape3, g)
    out = m.copy()
    for a, p in enumerate(shape3):
        if p <= 1:
            continue
        for idx in (g - 1, p - g):
            src = [slice(None)] * 3
            src[a] = g if idx == g - 1 else p - 1 - g
            dst = [slice(None)] * 3
            dst[a] = idx
            out[tuple(dst)] |= m[tuple(src)]
    return out
