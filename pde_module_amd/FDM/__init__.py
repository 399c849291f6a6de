"""Uniform-grid finite differences (reference ``pm/FDM/__init__.py:1-13``)."""
from .module import (Divergence, ExplicitUniformGridStencil, FarField, Grad, Laplacian,
                     OuterProduct, PeriodicLaplacian, ViscousDampingLayer, scalarVectorMult)
from .boundary import GridBoundary, ImmersedBoundary, PeriodicBoundary
from . import functional

__all__ = ["Laplacian", "PeriodicLaplacian", "Grad", "Divergence", "GridBoundary", "ImmersedBoundary", "PeriodicBoundary",
           "ExplicitUniformGridStencil", "FarField", "ViscousDampingLayer", "OuterProduct", "scalarVectorMult", "functional"]
