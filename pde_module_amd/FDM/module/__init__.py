from .explicitUniformGridStencil import ExplicitUniformGridStencil
from .laplacian import Laplacian
from .grad import Grad
from .divergence import Divergence

__all__ = ["ExplicitUniformGridStencil", "Laplacian", "Grad", "Divergence"]
