"""Small helpers (reference ``pm/utils/utils.py``, ``pm/utils/warp_utils.py``)."""
from __future__ import annotations


def get_unique_key(dictionary: dict, base_name: str) -> str:
    """base_name, or base_name_1, base_name_2, ... whichever is free first."""
    candidate, counter = base_name, 1
    while candidate in dictionary:
        candidate = f"{base_name}_{counter}"
        counter += 1
    return candidate


def ijk_to_global_c(i, j, k, Nx, Ny, Nz):
    """Flat C-order index of node (i, j, k)."""
    return i * (Ny * Nz) + j * Nz + k


def global_to_ijk_c(global_id, Nx, Ny, Nz):
    """(i, j, k) of a flat C-order node index."""
    return global_id // (Ny * Nz), (global_id // Nz) % Ny, global_id % Nz


class SignatureMismatchError(Exception):
    """Raised when two callables that must share a signature do not
    (reference ``pm/utils`` ``SignatureMismatchError``, raised by the
    Runge-Kutta integrators)."""
