"""ctypes binding of libpdehip.so (declared in include/pde_hip.h).

There is no CPU or PyTorch fallback: if the library is missing or no GPU is
visible, the first operator call raises.  Build the library with
``python -m pde_module_amd.build``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import torch

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libpdehip.so")

_P, _I64, _I32, _F = C.c_void_p, C.c_int64, C.c_int32, C.c_float

# name -> argtypes, in the order of include/pde_hip.h
SIGNATURES = {
    "pde_laplacian_f32": [_P, _P, _I64, _I64, _I64, _I32, _I32, _F, _F, _F, _F, _P],
    "pde_laplacian_generic_f32": [_P, _P, _I64, _I64, _I64, _I32, _I32, _F, _F, _F, _F, _P],
    "pde_laplacian_periodic_f32": [_P, _P, _I64, _I64, _I64, _I32, _I32, _F, _F, _F, _F, _P],
    "pde_grad_f32": [_P, _P, _I64, _I64, _I64, _I32, _I32, _F, _F, _F, _F, _P],
    "pde_divergence_f32": [_P, _P, _I64, _I64, _I64, _I32, _I32, _F, _F, _F, _F, _P],
    "pde_divergence_batch_f32": [_P, _P, _I64, _I64, _I64, _I64, _I32, _I32, _F, _F, _F, _F,
                                 _P],
    "pde_fused_lap_grad_div_f32": [_P, _P, _P, _P, _P, _I64, _I64, _I64, _I32,
                                   _F, _F, _F, _F, _F, _F, _F, _F, _F, _P],
    "pde_forward_euler_f32": [_P, _P, _P, _I64, _F, _P],
    "pde_copy_f32": [_P, _P, _I64, _P],
    "pde_heat_step_f32": [_P, _P, _I64, _I64, _I64, _I32, _I32, _F, _F, _F, _F, _F, _P],
    "pde_wave_step_f32": [_P, _P, _P, _P, _P, _I64, _I64, _I64, _I32, _F, _F, _F, _F, _F,
                          _I32, _P],
    "pde_bc_apply_f32": [_P, _P, _I64, _P, _I64, _P, _P],
    "pde_bc_apply_inplace_f32": [_P, _P, _I64, _P, _P, _P],
    "pde_periodic_fill_f32": [_P, _I64, _I64, _I64, _I32, _I32, _P],
    "pde_laplacian_adjoint_f32": [_P, _P, _I64, _I64, _I64, _I32, _I32, _F, _F, _F, _F, _P],
    "pde_grad_adjoint_f32": [_P, _P, _I64, _I64, _I64, _I32, _I32, _F, _F, _F, _F, _P],
    "pde_divergence_adjoint_f32": [_P, _P, _I64, _I64, _I64, _I32, _I32, _F, _F, _F, _F, _P],
    "pde_heat_step_adjoint_f32": [_P, _P, _I64, _I64, _I64, _I32, _I32, _F, _F, _F, _F, _F, _P],
    "pde_wave_step_adjoint_f32": [_P, _P, _P, _P, _P, _P, _P, _I64, _I64, _I64, _I32, _F, _F,
                                  _F, _F, _F, _I32, _P],
    "pde_bc_adjoint_f32": [_P, _P, _I64, _P, _I64, _P, _P, _P, _P, _P, _I64, _P],
    "pde_scale_f32": [_P, _P, _I64, _F, _P],
    "pde_bc_apply_batch_inplace_f32": [_P, _P, _I64, _P, _P, _P],
    "pde_batch_bc_capacity": [],
    "pde_bc_apply_inplace_noalloc_f32": [_P, _P, _I64, _P, _P],
    "pde_jacobi_neumann0_multisweep_f32": [_P, _P, _P, _I64, _I64, _I64, _I64, _I32, _F, _F,
                                           _I32, _P],
    "pde_jacobi_neumann0_adjoint_multisweep_f32": [_P, _P, _P, _P, _I64, _I64, _I64, _I64,
                                                   _I32, _F, _F, _I32, _I32, _P],
    "pde_heat_multistep_f32": [_P, _P, _P, _I64, _P, _I64, _P, _P, _P, _I32, _P],
    "pde_heat_batch_step_f32": [_P, _P, _P, _I64, _P, _I64, _I32, _I32, _I64, _P],
    "pde_outer_f32": [_P, _P, _P, _I64, _I32, _I32, _P],
    "pde_scalar_mul_f32": [_P, _P, _P, _I64, _I32, _P],
    "pde_div_scalar_f32": [_P, _P, _P, _I64, _I32, _P],
    "pde_matvec_f32": [_P, _P, _P, _I64, _I32, _I32, _P],
    "pde_lincomb_f32": [_P, _P, _P, _P, _F, _F, _F, _F, _I32, _P, _I64, _P],
    "pde_affine_f32": [_P, _P, _I64, _I32, _I32, _F, _F, _P],
    "pde_bc_adjoint_inplace_f32": [_P, _P, _I64, _P, _P, _P, _P, _P, _I64, _P, _P],
    "pde_bc_apply_chunked_inplace_f32": [_P, _P, _P, _I64, _P, _P],
    "pde_bc_apply_chunked_inplace_save_f32": [_P, _P, _P, _I64, _I64, _P, _P, _P, _P],
    "pde_heat_step_ring_f32": [_P, _P, _I64, _I64, _I64, _I32, _P, _I64, _P, _F, _F, _F, _F, _F,
                               _P, _P],
    "pde_wave_reverse_adjoint_pair_f32": [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I64, _I64,
                                          _I64, _I32, _F, _F, _F, _F, _F, _P, _P, _P, _P, _P,
                                          _I64, _P, _P],
    "pde_bc_adjoint_chunked_inplace_f32": [_P, _P, _P, _P, _P, _P, _P, _P, _P, _I64, _P],
    "pde_gather_f32": [_P, _P, _P, _I64, _P],
    "pde_wave_step_ctr_f32": [_P, _P, _P, _P, _P, _I64, _I64, _I64, _I32, _F, _F, _F, _F, _F,
                              _P, _P],
    "pde_wave_step_adjoint_ctr_f32": [_P, _P, _P, _P, _P, _P, _P, _I64, _I64, _I64, _I32, _F,
                                      _F, _F, _F, _F, _P, _P],
    "pde_gather_step_f32": [_P, _P, _P, _I64, _I64, _P, _P],
    "pde_scatter_add_step_f32": [_P, _P, _P, _I64, _I64, _F, _P, _P],
    "pde_counter_add": [_P, _I64, _P],
    "pde_counter_copy": [_P, _P, _I64, _P],
    "pde_wave_step_fused_f32": [_P, _P, _P, _P, _P, _I64, _I64, _I64, _I32, _F, _F, _F, _F, _F,
                                _P, _P, _P, _P, _P],
    "pde_wave_step_adjoint_fused_f32": [_P, _P, _P, _P, _P, _P, _P, _I64, _I64, _I64, _I32, _F,
                                        _F, _F, _F, _F, _P, _P, _P, _P],
    "pde_scatter_add_f32": [_P, _P, _P, _I64, _F, _P],
    "pde_damping_layer_f32": [_P, _P, _P, _I64, _P, _I32, _I32, _P, _F, _P],
    "pde_immersed_apply_f32": [_P, _P, _P, _I64, _P, _I64, _F, _I32, _I32, _I64, _I64, _I64,
                               _F, _P],
    "pde_jacobi_f32": [_P, _P, _P, _I64, _I64, _I64, _I32, _F, _F, _P],
    "pde_jacobi_neumann0_f32": [_P, _P, _P, _I64, _I64, _I64, _I32, _F, _F, _P],
    "pde_jacobi_adjoint_f32": [_P, _P, _P, _I64, _I64, _I64, _I32, _F, _F, _P],
    "pde_jacobi_neumann0_adjoint_f32": [_P, _P, _P, _I64, _I64, _I64, _I32, _F, _F, _P],
    "pde_jacobi_neumann0_batch_f32": [_P, _P, _P, _I64, _I64, _I64, _I64, _I32, _F, _F, _P],
    "pde_jacobi_neumann0_x2_f32": [_P, _P, _P, _I64, _I64, _I64, _I64, _I32, _F, _F, _P],
    "pde_cavity_predict_batch_f32": [_P, _P, _I64, _I64, _I64, _I64, _I32, _F, _F, _F, _F, _F,
                                     _F, _F, _F, _P],
    "pde_cavity_correct_batch_f32": [_P, _P, _P, _I64, _I64, _I64, _I64, _I32, _F, _F, _F, _F,
                                     _P],
    "pde_cavity_adjoint_correct_batch_f32": [_P, _P, _I64, _I64, _I64, _I64, _I32, _F, _F, _F,
                                             _F, _P],
    "pde_cavity_adjoint_div_batch_f32": [_P, _P, _P, _I64, _I64, _I64, _I64, _I32, _F, _F, _F,
                                         _F, _P],
    "pde_cavity_adjoint_predict_batch_f32": [_P, _P, _P, _I64, _I64, _I64, _I64, _I32, _F, _F,
                                             _F, _F, _F, _F, _F, _F, _P],
    "pde_cavity_predict_f32": [_P, _P, _I64, _I64, _I64, _I32, _F, _F, _F, _F, _F, _F, _F, _F,
                               _P],
    "pde_cavity_correct_f32": [_P, _P, _P, _I64, _I64, _I64, _I32, _F, _F, _F, _F, _P],
    "pde_cavity_adjoint_correct_f32": [_P, _P, _I64, _I64, _I64, _I32, _F, _F, _F, _F, _P],
    "pde_cavity_adjoint_div_f32": [_P, _P, _P, _I64, _I64, _I64, _I32, _F, _F, _F, _F, _P],
    "pde_cavity_adjoint_predict_f32": [_P, _P, _P, _I64, _I64, _I64, _I32, _F, _F, _F, _F, _F,
                                       _F, _F, _F, _P],
    "pde_jacobi_neumann0_adjoint_x2_f32": [_P, _P, _P, _I64, _I64, _I64, _I64, _I32, _F, _F, _P],
    "pde_jacobi_neumann0_adjoint_batch_f32": [_P, _P, _P, _I64, _I64, _I64, _I64, _I32, _F, _F,
                                              _P],
    "pde_matvec_t_f32": [_P, _P, _P, _I64, _I32, _I32, _P],
    "pde_version": [],
    "pde_status_string": [C.c_int],
}

# Must match struct pde_bc_op in include/pde_hip.h (32 bytes).  "value"
# holds the int32 bits of the slot for array-valued ops (the header's union);
# set_bc_slots / bc_slots convert.
BC_OP_DTYPE = np.dtype([("dst", "<i8"), ("src", "<i8"), ("kind", "<i2"), ("vkind", "<i2"),
                        ("value", "<f4"), ("coef", "<f4"), ("tmax", "<f4")])
BC_SET, BC_COPY, BC_GHOST_D, BC_GHOST_N = 0, 1, 2, 3
BCV_CONST, BCV_SINE, BCV_ARRAY = 0, 1, 2
assert BC_OP_DTYPE.itemsize == 32


def bc_slot_bits(slots) -> np.ndarray:
    """Slots of array-valued ops as the float32 bit patterns stored in "value"."""
    return np.ascontiguousarray(slots, dtype=np.int32).view(np.float32)


def bc_slots(ops) -> np.ndarray:
    """The slots of array-valued ops (value field read as int32 bits)."""
    return np.ascontiguousarray(ops["value"]).view(np.int32)

# struct pde_batch_grid
BATCH_GRID_DTYPE = np.dtype([("offset", "<i8"), ("op_begin", "<i8"), ("op_end", "<i8"),
                             ("p1", "<i4"), ("p2", "<i4"), ("ghost", "<i4"),
                             ("reserved0", "<i4"), ("wm", "<f4"), ("wc", "<f4"), ("wp", "<f4"),
                             ("alpha", "<f4"), ("dt", "<f4"), ("reserved1", "<f4")])
assert BATCH_GRID_DTYPE.itemsize == 64

# struct pde_heat_tile / pde_heat_tile_op (pde_heat_multistep_f32)
HEAT_TILE_DTYPE = np.dtype([("grid", "<i4"), ("r0", "<i4"), ("c0", "<i4"), ("rh", "<i4"),
                            ("rw", "<i4"), ("o_r0", "<i4"), ("o_r1", "<i4"), ("o_c0", "<i4"),
                            ("o_c1", "<i4"), ("op_beg", "<i4"), ("op_end", "<i4"),
                            ("reserved", "<i4")])
assert HEAT_TILE_DTYPE.itemsize == 48
HEAT_TILE_OP_DTYPE = np.dtype([("ldst", "<i4"), ("lsrc", "<i4"), ("op", "<i4"),
                               ("reserved", "<i4")])
assert HEAT_TILE_OP_DTYPE.itemsize == 16
HEAT_TILE_PITCH = HEAT_TILE_ROWS = 128
HEAT_TILE_MAXOPS = 2048
JACOBI_MAX_SWEEPS = 48        # PDE_JACOBI_MAX_SWEEPS

# struct pde_ring_piece / pde_ring_entry / pde_ring_plan
RING_PIECE_DTYPE = np.dtype([("r0", "<i4"), ("r1", "<i4"), ("c0", "<i4"), ("c1", "<i4"),
                             ("w_beg", "<i4"), ("w_end", "<i4"), ("t_beg", "<i4"),
                             ("t_end", "<i4"), ("z_beg", "<i4"), ("z_end", "<i4"),
                             ("reserved", "<i4", (2,))])
assert RING_PIECE_DTYPE.itemsize == 48
RING_ENTRY_DTYPE = np.dtype([("node", "<i4"), ("a", "<i4"), ("b", "<i4"), ("flags", "<i4")])
assert RING_ENTRY_DTYPE.itemsize == 16
RING_PLAN_DTYPE = np.dtype([("depth", "<i4"), ("npieces", "<i4"), ("nops", "<i8"),
                            ("pieces", "<u8"), ("w", "<u8"), ("t", "<u8"), ("z", "<u8"),
                            ("cdst", "<u8"), ("csign", "<u8")])
assert RING_PLAN_DTYPE.itemsize == 64

# struct pde_point_set (pointers as integers)
POINT_SET_DTYPE = np.dtype([("row_start", "<u8"), ("col", "<u8"), ("slot", "<u8"),
                            ("data", "<u8"), ("stride", "<i8"), ("scale", "<f4"),
                            ("reserved", "<i4")])
assert POINT_SET_DTYPE.itemsize == 48


# struct pde_layer_point
LAYER_POINT_DTYPE = np.dtype([("node", "<i8"), ("level", "<i4"), ("reserved", "<i4")])
assert LAYER_POINT_DTYPE.itemsize == 16
LAYER_MAXC = 16


# struct pde_ib_op
IB_OP_DTYPE = np.dtype([("node", "<i8"), ("value", "<f4"), ("comp", "<i4"), ("kind", "<i4"),
                        ("mask", "<i4")])
assert IB_OP_DTYPE.itemsize == 24


class PDEError(RuntimeError):
    """A libpdehip entry point returned a non-zero status."""


_lib = None


def load(path: str = LIB_PATH):
    """Load the library (once) and bind every entry point."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: the HIP kernels are not built "
            "(run `python -m pde_module_amd.build`); there is no fallback path")
    lib = C.CDLL(path)
    for name, argtypes in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = C.c_int
    lib.pde_status_string.restype = C.c_char_p
    _lib = lib
    return lib


def call(name: str, *args) -> None:
    """Call an entry point and raise PDEError on a non-zero status."""
    lib = load()
    st = getattr(lib, name)(*args)
    if st != 0:
        msg = lib.pde_status_string(st).decode()
        raise PDEError(f"{name} failed: {msg} (status {st})")


def stream_of(t: torch.Tensor) -> int:
    """Raw hipStream_t of torch's current stream on t's device."""
    return torch.cuda.current_stream(t.device).cuda_stream


def dev_ptr(t: torch.Tensor) -> int:
    return t.data_ptr()


def check_field(t: torch.Tensor, name: str = "field") -> None:
    """Fields handed to the kernels: contiguous FP32 on a GPU."""
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor, got {type(t).__name__}")
    if not t.is_cuda:
        raise RuntimeError(f"{name} must live on a GPU (got device {t.device}); "
                           "the HIP kernels have no CPU path")
    if t.dtype != torch.float32:
        raise TypeError(f"{name} must be float32, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
