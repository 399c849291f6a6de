from .stencil import Stencil, tuplify
from .hooks import setup, before_forward, after_forward
from .elementWise import ElementWise
from .mapWise import MapWise
from . import element_ops

__all__ = ["Stencil", "tuplify", "setup", "before_forward", "after_forward", "ElementWise",
           "MapWise", "element_ops"]
