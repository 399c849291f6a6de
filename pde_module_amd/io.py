"""Field snapshots, VTK output and the Ghia cavity reference data (the field
I/O around the reference's examples: ``wp.array.numpy()`` snapshots,
``to_pyvista`` in ``pm/mesh/mesh_functions.py:11-20``, and the Ghia et al.
centreline tables read in ``examples/LDC/LDC_No_Graph.py:116-135``).

``write_vtk`` writes a legacy-format VTK file (``STRUCTURED_POINTS``: origin,
spacing and the padded node counts of a ``UniformGridMesh``) with one point
array per field, so results open in ParaView/VisIt without pyvista.
``to_pyvista`` returns a ``pyvista.ImageData`` of the same grid when pyvista
is installed.
"""
from __future__ import annotations

import csv
from typing import Mapping

import numpy as np
import torch


def to_numpy(field) -> np.ndarray:
    """Host copy of a field (the counterpart of ``wp.array.numpy()``)."""
    if isinstance(field, torch.Tensor):
        return field.detach().cpu().numpy()
    return np.asarray(field)


def _grid_geometry(grid):
    dims = tuple(int(n) for n in grid.nodes_per_axis)
    origin = tuple(float(v[0]) for v in grid.coordinate_vectors)
    spacing = tuple(float(grid.dx) for _ in range(3))
    return dims, origin, spacing


def write_vtk(path: str, grid, point_data: Mapping[str, object], binary: bool = True) -> None:
    """Legacy VTK file of the grid's nodes (ghosts included) with the given
    fields as point data: element size 1 as SCALARS, 3 as VECTORS (2-vectors
    padded with a zero z), anything else as a FIELD array of that many
    components.  Fields are (X, Y, Z) + element shape, as the modules use."""
    dims, origin, spacing = _grid_geometry(grid)
    npts = int(np.prod(dims))
    arrays = {}
    for name, f in point_data.items():
        a = np.asarray(to_numpy(f), dtype=np.float32)
        if tuple(a.shape[:3]) != dims:
            raise ValueError(f"field {name!r} has grid {a.shape[:3]}, the mesh has {dims}")
        # VTK orders points x fastest: transpose (X, Y, Z, C) to (Z, Y, X, C)
        comps = int(np.prod(a.shape[3:])) if a.ndim > 3 else 1
        a = a.reshape(dims + (comps,)).transpose(2, 1, 0, 3).reshape(npts, comps)
        if comps == 2:
            a = np.concatenate([a, np.zeros((npts, 1), np.float32)], axis=1)
            comps = 3
        arrays[name] = (a, comps)
    with open(path, "wb") as fh:
        w = lambda s: fh.write(s.encode("ascii"))
        w("# vtk DataFile Version 3.0\n")
        w("pde_module_amd field output\n")
        w("BINARY\n" if binary else "ASCII\n")
        w("DATASET STRUCTURED_POINTS\n")
        w(f"DIMENSIONS {dims[0]} {dims[1]} {dims[2]}\n")
        w(f"ORIGIN {origin[0]!r} {origin[1]!r} {origin[2]!r}\n")
        w(f"SPACING {spacing[0]!r} {spacing[1]!r} {spacing[2]!r}\n")
        w(f"POINT_DATA {npts}\n")

        def body(a):
            if binary:
                fh.write(a.astype(">f4").tobytes())
                w("\n")
            else:
                np.savetxt(fh, a, fmt="%.9g")

        others = []
        for name, (a, comps) in arrays.items():
            if comps == 1:
                w(f"SCALARS {name} float 1\nLOOKUP_TABLE default\n")
                body(a)
            elif comps == 3:
                w(f"VECTORS {name} float\n")
                body(a)
            else:
                others.append((name, a, comps))
        if others:
            w(f"FIELD FieldData {len(others)}\n")
            for name, a, comps in others:
                w(f"{name} {comps} {npts} float\n")
                body(a)


def read_vtk_structured_points(path: str) -> dict:
    """Reads back what ``write_vtk`` writes: {'dims', 'origin', 'spacing',
    name: (npts, comps) array}."""
    out = {}
    with open(path, "rb") as fh:
        data = fh.read()
    pos = 0

    def line():
        nonlocal pos
        end = data.index(b"\n", pos)
        s = data[pos:end].decode("ascii")
        pos = end + 1
        return s

    line(), line()
    binary = line().strip() == "BINARY"
    line()
    out["dims"] = tuple(int(v) for v in line().split()[1:])
    out["origin"] = tuple(float(v) for v in line().split()[1:])
    out["spacing"] = tuple(float(v) for v in line().split()[1:])
    npts = int(line().split()[1])

    def values(count):
        nonlocal pos
        if binary:
            a = np.frombuffer(data[pos:pos + 4 * count], dtype=">f4").astype(np.float32)
            pos += 4 * count + 1
            return a
        vals = []
        while len(vals) < count:
            vals += [float(v) for v in line().split()]
        return np.array(vals, np.float32)

    while pos < len(data):
        head = line().split()
        if not head:
            continue
        if head[0] == "SCALARS":
            line()
            out[head[1]] = values(npts).reshape(npts, 1)
        elif head[0] == "VECTORS":
            out[head[1]] = values(3 * npts).reshape(npts, 3)
        elif head[0] == "FIELD":
            for _ in range(int(head[2])):
                name, comps, n, _t = line().split()
                out[name] = values(int(comps) * int(n)).reshape(int(n), int(comps))
    return out


def to_pyvista(grid):
    """``pyvista.ImageData`` of the grid's nodes (``to_pyvista`` in the
    reference builds an unstructured grid from the mesh cells)."""
    try:
        import pyvista as pv
    except ImportError as e:
        raise ImportError("pyvista is not installed; write_vtk writes the same grid "
                          "to a legacy VTK file") from e
    dims, origin, spacing = _grid_geometry(grid)
    return pv.ImageData(dimensions=dims, origin=origin, spacing=spacing)


def read_ghia(path: str) -> dict:
    """Ghia, Ghia & Shin (1982) centreline table as the reference's examples
    ship it (tab- or comma-separated, a '%x' / '%y' coordinate column, then
    one column per Reynolds number): {column name: float array}."""
    with open(path, newline="") as fh:
        text = fh.read()
    delim = "\t" if "\t" in text.splitlines()[0] else ","
    rows = list(csv.reader(text.splitlines(), delimiter=delim))
    head = [h.strip() for h in rows[0]]
    cols = {h: [] for h in head}
    for r in rows[1:]:
        if not r or not r[0].strip():
            continue
        for h, v in zip(head, r):
            cols[h].append(float(v))
    return {h: np.array(v) for h, v in cols.items()}


def centreline(field2d: np.ndarray, coords: np.ndarray, along: int, at: float, comp: int,
               sample: np.ndarray) -> np.ndarray:
    """Linear interpolation of component ``comp`` of a 2D field (X, Y, C) on
    the grid line ``coords[other axis] == at`` (interpolated between the two
    nearest lines) at the positions ``sample`` along axis ``along``."""
    other = 1 - along
    c_other = coords[other]
    k = int(np.searchsorted(c_other, at))
    k = min(max(k, 1), len(c_other) - 1)
    t = (at - c_other[k - 1]) / (c_other[k] - c_other[k - 1])
    if along == 0:
        line = (1 - t) * field2d[:, k - 1, comp] + t * field2d[:, k, comp]
    else:
        line = (1 - t) * field2d[k - 1, :, comp] + t * field2d[k, :, comp]
    return np.interp(sample, coords[along], line)
