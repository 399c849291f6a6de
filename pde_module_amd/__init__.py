"""pde_module_amd — MI355X-native uniform-grid finite differences.

A drop-in for PDE_Module's Laplacian / Grad / Divergence modules, its ghost-cell
GridBoundary (functional and legacy) and ForwardEuler, running hand-written
gfx950 HIP kernels from libpdehip.so through a C ABI (include/pde_hip.h).
Fields are FP32 torch tensors on the GPU of padded shape (X, Y, Z) + element
shape (C,) or (N, D).
"""
from . import _native
from .FDM import (Divergence, Grad, GridBoundary, Laplacian, PeriodicBoundary,
                  PeriodicLaplacian)
from .time_step import ForwardEuler, HeatStep, WaveStep
from .mesh import UniformGridMesh, create_structured_field

__all__ = ["Laplacian", "PeriodicLaplacian", "Grad", "Divergence", "GridBoundary", "PeriodicBoundary",
           "ForwardEuler", "HeatStep", "WaveStep", "UniformGridMesh",
           "create_structured_field"]
