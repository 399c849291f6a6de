"""Module-level host behaviour that needs no GPU: lifecycle hooks, shapes and
the loud failure when a field is not on a GPU (there is no CPU fallback)."""
import numpy as np
import pytest
import torch

from pde_module_amd import ForwardEuler, Grad, Divergence, Laplacian, UniformGridMesh
from pde_module_amd.mesh import create_structured_field
from pde_module_amd.stencil import Stencil, after_forward, before_forward, setup


def test_hooks_run_in_order():
    calls = []

    class S(Stencil):
        @setup(order=2)
        def b(self, x):
            calls.append("setup2")

        @setup(order=-1)
        def a(self, x):
            calls.append("setup-1")

        @before_forward
        def c(self, x):
            calls.append("before")

        @after_forward
        def d(self, x):
            calls.append("after")

        def forward(self, x):
            calls.append("forward")
            return x

    s = S()
    s(1)
    s(2)
    assert calls == ["setup-1", "setup2", "before", "forward", "after", "before", "forward", "after"]


def test_double_registration_rejected():
    with pytest.raises(AssertionError):
        @setup
        @before_forward
        def f(self):
            pass


def test_cpu_tensor_fails_loudly():
    u = torch.zeros(6, 6, 1, 1)
    lap = Laplacian(1, 0.1, 1)
    with pytest.raises(RuntimeError, match="GPU"):
        lap(u, 1.0)
    e = ForwardEuler()
    with pytest.raises(RuntimeError, match="GPU"):
        e(u, u, 0.1)


def test_shapes_and_weights():
    lap = Laplacian(3, 0.1, 1)
    assert lap.input_dtype == (3,) and lap.output_dtype == (3,)
    assert lap.stencil == tuple(np.float32(v) for v in (100.0, -200.0, 100.0))
    g = Grad(1, (10, 10, 1), 0.1, 1)
    assert g.output_dtype == (2,)
    assert Grad(1, (10, 10, 1), 0.1, 1, force_matrix=True).output_dtype == (1, 2)
    assert Grad(2, (10, 10, 10), 0.1, 1).output_dtype == (2, 3)
    assert Divergence(2, (10, 10, 1), 0.1, 1).output_dtype == (1,)
    assert Divergence((2, 2), (10, 10, 1), 0.1, 1).output_dtype == (2,)
    with pytest.raises(AssertionError):
        Divergence(3, (10, 10, 1), 0.1, 1)
    with pytest.raises(ValueError):
        Laplacian(1, 0.1, 1, stencil=(1, -2, 1))


def test_mesh_layout_contract():
    m = UniformGridMesh(0.25, (5, 3, 1), ghost_cells=1)
    assert m.grid_shape == (7, 5, 1)
    np.testing.assert_allclose(m.coordinate_vectors[0], np.arange(-1, 6) * 0.25)
    assert m.meshgrid[0].shape == (7, 5, 1)
    f = create_structured_field(m, "node", 2, device="cpu")
    assert tuple(f.shape) == (7, 5, 1, 2)
    f = create_structured_field(m, "node", 1, func=lambda x, y, z: x + 10 * y, device="cpu")
    assert float(f[2, 3, 0, 0]) == pytest.approx(0.25 + 10 * 0.5)


def test_grid_shape_may_carry_element_dims():
    """Grad/Divergence take the field's shape; a torch shape ends with the
    element dims, which are not grid axes."""
    Grad(2, (23, 23, 1, 2), 0.05, 1)
    Divergence(2, (23, 23, 1, 2), 0.05, 1)
    with pytest.raises(AssertionError):
        Divergence(3, (23, 23, 1, 3), 0.05, 1)


def test_element_ops_take_functions_and_refuse_unknown_names():
    """Compiled ops by name; any other function runs as device torch
    operations; unknown names and non-callables raise; CPU fields are refused
    (no host fallback) for functions too."""
    from pde_module_amd.stencil import ElementWise, MapWise, element_ops
    fe = ElementWise(lambda a, b: a + b)
    assert isinstance(fe.element_op, element_ops.FunctionElementOp)
    assert isinstance(MapWise(lambda m, rho: m / rho[..., 0:1]).element_op,
                      element_ops.FunctionMap)
    with pytest.raises(TypeError):
        MapWise("not_a_map")
    with pytest.raises(TypeError):
        ElementWise(3.0)
    assert ElementWise("outer").element_op is element_ops.outer
    x = torch.zeros(4, 4, 1, 2)
    with pytest.raises(RuntimeError):
        ElementWise("add")(x, x)
    with pytest.raises(RuntimeError):
        fe(x, x)
