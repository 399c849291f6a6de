"""Host resolution of GridBoundary groups into the device op table (CPU only).

The resolved table is interpreted on the host (tests/helpers.py) and compared
bit for bit with the oracle's procedural restatement of the reference's copy +
sequential group launches."""
import numpy as np
import pytest
import torch

from conftest import same_bits
from helpers import apply_ops_host
from oracle import np_oracle as O
from pde_module_amd.deprecated import GridBoundary as LegacyGridBoundary
from pde_module_amd.FDM import GridBoundary
from pde_module_amd.FDM.boundary import get_ramp_func


def field(shape3, C):
    return torch.zeros(shape3 + (C,), dtype=torch.float32)


def groups_for_oracle(bc):
    out = []
    for fb in bc.func_groups.values():
        comps = fb.output_ids if isinstance(fb.output_ids, (list, tuple)) else [fb.output_ids]
        tmax = getattr(fb.func, "T_max", None)
        out.append(dict(ids=fb.face_ids, type=fb.boundary_type, comps=list(comps),
                        const=fb.func.constant, tmax=tmax))
    return out


FUNCTIONAL_CASES = {
    "dirichlet_all": ((12, 10, 1), 1, lambda b: b.dirichlet_BC("ALL", 0.0)),
    "neumann_all": ((9, 11, 1), 1, lambda b: b.vonNeumann_BC("ALL", 0.37)),
    "lid_cavity": ((10, 10, 1), 2, lambda b: (b.no_slip("ALL"), b.dirichlet_BC("+Y", 1.0, 0))),
    "impermeable": ((8, 9, 1), 2, lambda b: (b.impermeable("-X"), b.impermeable("+Y"),
                                              b.dirichlet_BC("+X", 0.5))),
    "mixed_3d": ((7, 6, 8), 3, lambda b: (b.dirichlet_BC("ALL", 1.25, [0, 2]),
                                          b.vonNeumann_BC("-Z", -2.0, 1),
                                          b.dirichlet_BC("+X", 0.0))),
    "ramp": ((9, 9, 1), 2, lambda b: (b.no_slip("ALL"),
                                      b.dirichlet_BC("+Y", get_ramp_func(1.0, 0.5), 0))),
    "explicit_ids": ((8, 8, 1), 1, lambda b: b.dirichlet_BC([9, 10, 27, 35], 3.0)),
    "ghost2": ((10, 9, 1), 1, lambda b: (b.vonNeumann_BC("-Y", 1.0), b.dirichlet_BC("+X", 2.0))),
}


@pytest.mark.parametrize("name", sorted(FUNCTIONAL_CASES))
@pytest.mark.parametrize("t", [0.0, 0.2, 0.9])
def test_functional_table_matches_oracle(corc, name, t):
    shape3, C, configure = FUNCTIONAL_CASES[name]
    g = 2 if name == "ghost2" else 1
    bc = GridBoundary(field(shape3, C), 0.1, g)
    configure(bc)
    ops = bc.resolved_ops()
    assert len(np.unique(ops["dst"])) == len(ops)
    rng = np.random.default_rng(len(name))
    u = rng.uniform(-1, 1, shape3 + (C,)).astype(np.float32)
    want = corc.bc_functional(u, g, groups_for_oracle(bc), 0.1, t=t)
    got = apply_ops_host(u, ops, t)
    assert same_bits(got, want)


def test_functional_groups_match_oracle():
    for shape3, g in [((9, 8, 1), 1), ((6, 7, 8), 2), ((1, 7, 9), 1)]:
        bc = GridBoundary(field(shape3, 1), 0.1, g)
        want = O.functional_groups(shape3, g)
        assert set(bc.groups) == set(want)
        for k in want:
            np.testing.assert_array_equal(np.asarray(bc.groups[k]), want[k])


LEGACY_CASES = {
    "default": ((10, 12, 1), 1, lambda b: None),
    "dirichlet_all": ((10, 12, 1), 1, lambda b: b.dirichlet_BC("ALL", 0.0)),
    "lid": ((9, 9, 1), 2, lambda b: (b.no_slip("ALL"), b.dirichlet_BC("+Y", 1.0, 0))),
    "neumann": ((8, 10, 1), 1, lambda b: b.vonNeumann_BC("-X", 0.5)),
    "impermeable_3d": ((6, 7, 8), 3, lambda b: (b.impermeable("-Z"), b.dirichlet_BC("+X", 2.0, [1]))),
    "shifted": ((12, 10, 1), 1, lambda b: (b.vonNeumann_BC("-X", 0.25), b.shift_group("-X", 2, 0))),
}


@pytest.mark.parametrize("name", sorted(LEGACY_CASES))
def test_legacy_table_matches_oracle(corc, name):
    shape3, C, configure = LEGACY_CASES[name]
    bc = LegacyGridBoundary(field(shape3, C), 0.1, 1)
    configure(bc)
    ops = bc.resolved_ops()
    rng = np.random.default_rng(len(name))
    u = rng.uniform(-1, 1, shape3 + (C,)).astype(np.float32)
    want = corc.bc_legacy(u, bc.boundary_ijk_indices, bc.boundary_interior,
                          bc.boundary_type, bc.boundary_value, 0.1)
    assert same_bits(apply_ops_host(u, ops), want)


def test_legacy_layout_matches_oracle():
    for shape3 in [(9, 8, 1), (6, 7, 8)]:
        bc = LegacyGridBoundary(field(shape3, 1), 0.1, 1)
        ijk, groups, inter = O.legacy_boundary_layout(shape3, 1)
        np.testing.assert_array_equal(bc.boundary_ijk_indices, ijk)
        np.testing.assert_array_equal(bc.boundary_interior, inter)
        for k in groups:
            np.testing.assert_array_equal(bc.groups[k], groups[k])


def test_bad_arguments_raise():
    bc = GridBoundary(field((6, 6, 1), 2), 0.1, 1)
    with pytest.raises(AssertionError):
        bc.dirichlet_BC("+W", 0.0)
    with pytest.raises(TypeError):
        bc.dirichlet_BC("ALL", lambda *a: 0.0)
    with pytest.raises(AssertionError):
        bc.dirichlet_BC("ALL", 0.0, 5)
    with pytest.raises(AssertionError):
        GridBoundary(field((6, 6, 1), 3), 0.1, 1).impermeable("-X")
    leg = LegacyGridBoundary(field((6, 6, 1), 1), 0.1, 1)
    with pytest.raises(NotImplementedError):
        leg.dirichlet_BC("ALL", lambda *a: 0.0)
    with pytest.raises(AssertionError):
        LegacyGridBoundary(field((6, 6, 1), 1), 0.1, 0)


def test_out_of_grid_tables_raise_on_host():
    from pde_module_amd.FDM.boundary.device_table import DeviceBoundaryTable
    bc = GridBoundary(field((6, 6, 1), 1), 0.1, 1)
    with pytest.raises(IndexError):
        bc.dirichlet_BC([36], 0.0)
    leg = LegacyGridBoundary(field((8, 8, 1), 1), 0.1, 1)
    leg.shift_group("+X", 2, 0)           # +X nodes move onto the last ghost layer
    ops = leg.resolved_ops()
    tab = DeviceBoundaryTable.__new__(DeviceBoundaryTable)
    src = ops["src"][ops["kind"] != 0]
    tab.max_index = int(max(ops["dst"].max(), src.max()))
    tab.min_index = int(min(ops["dst"].min(), src.min()))
    with pytest.raises(IndexError):
        tab._check_bounds(64)
