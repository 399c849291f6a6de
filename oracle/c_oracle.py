"""ORACLE — test infrastructure only (tests/, smoke() and bench.py's cpu_baseline).

ctypes front end of the C FP32 restatement in ``pde_oracle.c``.  All arrays are
NumPy float32/int arrays on the host; the functions mirror the NumPy oracle's
signatures so tests can run both on the same inputs.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libpdeoracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_i8p = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")
_I64, _I, _F = C.c_int64, C.c_int, C.c_float


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_laplacian_f32.argtypes = [_f32p, _f32p, _I64, _I64, _I64, _I, _I, _F, _F, _F, _F, _I]
        L.orc_grad_f32.argtypes = [_f32p, _f32p, _I64, _I64, _I64, _I, _I, _F, _F, _F, _F, _I]
        L.orc_divergence_f32.argtypes = [_f32p, _f32p, _I64, _I64, _I64, _I, _I, _I, _F, _F, _F, _F, _I]
        L.orc_forward_euler_f32.argtypes = [_f32p, _f32p, _f32p, _I64, _F, _I]
        L.orc_bc_legacy_f32.argtypes = [_f32p, _f32p, _I64, _I64, _I64, _I, _i32p, _i32p, _i8p, _f32p, _I64, _F]
        L.orc_bc_functional_f32.argtypes = [_f32p, _f32p, _I64, _I64, _I64, _I, _I, _I, _i64p, _i32p,
                                            _i64p, _i32p, _i8p, _f32p, _f32p, _F, _F]
        L.orc_periodic_fill_f32.argtypes = [_f32p, _I64, _I64, _I64, _I, _I]
        L.orc_wave_step_f32.argtypes = [_f32p, _f32p, C.c_void_p, _f32p, _f32p, _I64, _I64, _I64,
                                        _I, _F, _F, _F, _F, _F, _I, _I]
        L.orc_max_threads.restype = C.c_int
        _lib = L
    return _lib


def max_threads():
    return lib().orc_max_threads()


def _nan_like(u, shape=None):
    return np.full(u.shape if shape is None else shape, np.nan, dtype=np.float32)


def laplacian(u, w, alpha, g, nthreads=1, out=None):
    X, Y, Z, Cc = u.shape
    out = _nan_like(u) if out is None else out
    lib().orc_laplacian_f32(u, out, X, Y, Z, Cc, g, w[0], w[1], w[2], alpha, nthreads)
    return out


def grad(u, w, alpha, g, force_matrix=False, nthreads=1):
    X, Y, Z, N = u.shape
    D = sum(1 for s in (X, Y, Z) if s > 1)
    out = _nan_like(u, (X, Y, Z, N, D))
    lib().orc_grad_f32(u, out, X, Y, Z, N, g, w[0], w[1], w[2], alpha, nthreads)
    if N == 1 and not force_matrix:
        return out.reshape(X, Y, Z, D)
    return out


def divergence(u, w, alpha, g, matrix_in=False, nthreads=1):
    X, Y, Z = u.shape[:3]
    rows = u.shape[3] if matrix_in else 1
    u = np.ascontiguousarray(u)
    out = _nan_like(u, (X, Y, Z, rows))
    lib().orc_divergence_f32(u, out, X, Y, Z, rows, int(matrix_in), g, w[0], w[1], w[2], alpha, nthreads)
    return out


def forward_euler(cur, k, dt_step, nthreads=1):
    out = np.empty_like(cur)
    lib().orc_forward_euler_f32(cur.ravel(), np.ascontiguousarray(k).ravel(), out.ravel(),
                                cur.size, dt_step, nthreads)
    return out


def bc_legacy(u, ijk, inter, btype, bvalue, dx):
    X, Y, Z, Cc = u.shape
    out = np.empty_like(u)
    lib().orc_bc_legacy_f32(u, out, X, Y, Z, Cc, np.ascontiguousarray(ijk, np.int32),
                            np.ascontiguousarray(inter, np.int32), np.ascontiguousarray(btype, np.int8),
                            np.ascontiguousarray(bvalue, np.float32), len(ijk), np.float32(dx))
    return out


def bc_functional(u, g, groups, dx, t=0.0):
    X, Y, Z, Cc = u.shape
    out = np.empty_like(u)
    off = np.zeros(len(groups) + 1, np.int64)
    coff = np.zeros(len(groups) + 1, np.int64)
    for i, grp in enumerate(groups):
        off[i + 1] = off[i] + len(grp["ids"])
        coff[i + 1] = coff[i] + len(grp["comps"])
    ids = np.concatenate([np.asarray(x["ids"], np.int32) for x in groups]) if groups else np.zeros(1, np.int32)
    comps = np.concatenate([np.asarray(x["comps"], np.int32) for x in groups]) if groups else np.zeros(1, np.int32)
    gtype = np.array([x["type"] for x in groups] or [0], np.int8)
    gconst = np.array([x["const"] for x in groups] or [0], np.float32)
    gtmax = np.array([(x.get("tmax") or 0.0) for x in groups] or [0], np.float32)
    lib().orc_bc_functional_f32(u, out, X, Y, Z, Cc, g, len(groups), off, ids, coff, comps,
                                gtype, gconst, gtmax, np.float32(t), np.float32(dx))
    return out


def periodic_fill(u, g):
    X, Y, Z, Cc = u.shape
    lib().orc_periodic_fill_f32(u, X, Y, Z, Cc, g)
    return u


def wave_step_core(u2, v, c2, w, alpha, g, dt_step, first, nthreads=1):
    """np_oracle.wave_step_core in C: (un, vn) of the fused u/v wave update on
    the boundary-applied u2 (scalar fields; c2 None means 1)."""
    X, Y, Z, Cc = u2.shape
    if Cc != 1:
        raise ValueError("the C wave step takes scalar fields")
    un, vn = np.empty_like(u2), np.empty_like(v)
    c2p = None if c2 is None else np.ascontiguousarray(c2, dtype=np.float32)
    lib().orc_wave_step_f32(u2, v, None if c2p is None else c2p.ctypes.data_as(C.c_void_p),
                            un, vn, X, Y, Z, g, w[0], w[1], w[2], np.float32(alpha),
                            np.float32(dt_step), int(bool(first)), nthreads)
    return un, vn
