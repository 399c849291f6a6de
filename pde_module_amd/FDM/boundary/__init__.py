from .flags import DIRICHLET, VON_NEUMANN, NO_BC, Boundary_Types
from .boundary_functions import get_constant_func, get_ramp_func, get_sine_func
from .functional_boundary import GridBoundary
from .periodic import PeriodicBoundary
from .immersedBoundary import ImmersedBoundary

__all__ = ["GridBoundary", "PeriodicBoundary", "ImmersedBoundary", "DIRICHLET", "VON_NEUMANN", "NO_BC",
           "Boundary_Types", "get_constant_func", "get_ramp_func", "get_sine_func"]
