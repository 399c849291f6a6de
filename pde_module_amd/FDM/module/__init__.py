from .explicitUniformGridStencil import ExplicitUniformGridStencil
from .laplacian import Laplacian
from .grad import Grad
from .divergence import Divergence
from .elementWiseOps import OuterProduct, scalarVectorMult

__all__ = ["ExplicitUniformGridStencil", "Laplacian", "Grad", "Divergence", "OuterProduct",
           "scalarVectorMult"]
