"""Boundary-condition type flags (reference ``pm/FDM/boundary/flags.py:8-21``,
``pm/utils/constants.py:10-25``)."""
from enum import IntEnum

NO_BC = 0
DIRICHLET = 1
VON_NEUMANN = 2
BC_SRING = {1: "DIRICHLET", 2: "VON_NEUMANN"}

FLUID_CELL = 0
SOLID_CELL = 1


class Boundary_Types(IntEnum):
    NO_BC = NO_BC
    DIRICHLET = DIRICHLET
    VON_NEUMANN = VON_NEUMANN
