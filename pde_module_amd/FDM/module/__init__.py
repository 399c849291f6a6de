from .explicitUniformGridStencil import ExplicitUniformGridStencil
from .laplacian import Laplacian, PeriodicLaplacian
from .grad import Grad
from .divergence import Divergence
from .elementWiseOps import OuterProduct, scalarVectorMult
from .farfield import FarField
from .viscousDampingLayer import ViscousDampingLayer

__all__ = ["ExplicitUniformGridStencil", "Laplacian", "PeriodicLaplacian", "Grad", "Divergence", "OuterProduct", "FarField", "ViscousDampingLayer",
           "scalarVectorMult"]
