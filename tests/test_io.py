"""Field I/O: VTK files round-trip (binary and ASCII, scalar / 2- / 3-vector /
matrix point data, x-fastest point order), the Ghia tables read from the
committed fixtures, centreline interpolation exact on linear fields (CPU)."""
import os

import numpy as np
import pytest

from pde_module_amd.io import centreline, read_ghia, read_vtk_structured_points, write_vtk
from pde_module_amd.mesh import UniformGridMesh

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("binary", [True, False])
def test_vtk_round_trip(tmp_path, binary):
    grid = UniformGridMesh(dx=0.25, nodes_per_axis=(4, 3, 1), origin=(0.0, 0.0, 0.0), ghost_cells=1)
    dims = grid.nodes_per_axis
    rng = np.random.default_rng(0)
    s = rng.uniform(-1, 1, dims + (1,)).astype(np.float32)
    v2 = rng.uniform(-1, 1, dims + (2,)).astype(np.float32)
    v3 = rng.uniform(-1, 1, dims + (3,)).astype(np.float32)
    m = rng.uniform(-1, 1, dims + (2, 2)).astype(np.float32)
    path = str(tmp_path / "f.vtk")
    write_vtk(path, grid, {"s": s, "u": v2, "w": v3, "J": m}, binary=binary)
    r = read_vtk_structured_points(path)
    assert r["dims"] == dims and r["spacing"] == (0.25, 0.25, 0.25)
    assert r["origin"] == (-0.25, -0.25, 0.0)
    # point p = i + X*(j + Y*k) holds node (i, j, k)
    i, j, k = 2, 1, 0
    p = i + dims[0] * (j + dims[1] * k)
    assert r["s"][p, 0] == s[i, j, k, 0]
    assert np.array_equal(r["u"][p], [v2[i, j, k, 0], v2[i, j, k, 1], 0.0])
    assert np.array_equal(r["w"][p], v3[i, j, k])
    assert np.array_equal(r["J"][p], m[i, j, k].reshape(-1))
    with pytest.raises(ValueError):
        write_vtk(path, grid, {"bad": np.zeros((3, 3, 1, 1), np.float32)})


def test_ghia_tables():
    u = read_ghia(os.path.join(HERE, "golden", "ghia_u_velocity_re.tsv"))
    v = read_ghia(os.path.join(HERE, "golden", "ghia_v_velocity_re.csv"))
    assert list(u)[:2] == ["%y", "100"] and list(v)[:2] == ["%x", "100"]
    assert len(u["%y"]) == 17 and len(v["%x"]) == 17
    assert u["100"][1] == 0.84123 and u["%y"][1] == 0.9766
    assert v["100"][1] == -0.05906 and v["%x"][1] == 0.9688
    assert u["100"][0] == 1.0 and v["100"][0] == 0.0


def test_centreline_exact_on_linear_fields():
    x = np.linspace(0, 1, 11)
    y = np.linspace(0, 1, 9)
    X, Y = np.meshgrid(x, y, indexing="ij")
    f = np.stack([2 * X + 3 * Y, -X + Y], axis=-1)
    s = np.array([0.1, 0.37, 0.9])
    assert np.allclose(centreline(f, (x, y), along=1, at=0.5, comp=0, sample=s), 2 * 0.5 + 3 * s)
    assert np.allclose(centreline(f, (x, y), along=0, at=0.33, comp=1, sample=s), -s + 0.33)
