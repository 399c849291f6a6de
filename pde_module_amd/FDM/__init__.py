"""Uniform-grid finite differences (reference ``pm/FDM/__init__.py:1-13``)."""
from .module import (Divergence, ExplicitUniformGridStencil, Grad, Laplacian, OuterProduct,
                     scalarVectorMult)
from .boundary import GridBoundary, PeriodicBoundary
from . import functional

__all__ = ["Laplacian", "Grad", "Divergence", "GridBoundary", "PeriodicBoundary",
           "ExplicitUniformGridStencil", "OuterProduct", "scalarVectorMult", "functional"]
