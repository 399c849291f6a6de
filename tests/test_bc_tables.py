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
        bc.dirichlet_BC("ALL", "not a value")
    bc.dirichlet_BC("ALL", lambda *a: 0.0)          # a boundary function: accepted
    with pytest.raises(AssertionError):
        bc.dirichlet_BC("ALL", 0.0, 5)
    with pytest.raises(AssertionError):
        GridBoundary(field((6, 6, 1), 3), 0.1, 1).impermeable("-X")
    leg = LegacyGridBoundary(field((6, 6, 1), 1), 0.1, 1)
    with pytest.raises(AssertionError):             # one component per function group
        leg.dirichlet_BC("ALL", lambda *a: 0.0)
    leg.dirichlet_BC("ALL", lambda *a: 0.0, 0)
    with pytest.raises(AssertionError):
        leg.dirichlet_BC("ALL", lambda *a: 0.0, 1)
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


def _chunk_cases():
    cases = []
    u = field((9, 8, 1), 1)
    b = LegacyGridBoundary(u, 0.1, 1)
    cases.append(("legacy2d", u, b))
    u = field((7, 8, 9), 3)
    b = LegacyGridBoundary(u, 0.1, 1)
    b.impermeable("+Z")
    b.vonNeumann_BC("-X", 0.5)
    cases.append(("legacy3d", u, b))
    u = field((30, 28, 1), 2)
    b = GridBoundary(u, 0.1, 1)
    b.no_slip("ALL")
    b.dirichlet_BC("+Y", 1.0, 0)
    cases.append(("lid", u, b))
    u = field((300, 200, 1), 1)
    b = LegacyGridBoundary(u, 0.1, 1)
    cases.append(("large", u, b))
    return cases


@pytest.mark.parametrize("case", range(4))
def test_chunked_tables_are_conflict_closed_and_exact(case):
    """Chunked in-place application (each chunk reads, then writes, chunks one
    after the other) equals the two-phase apply; the chunked transpose applied
    the same way equals the transpose oracle."""
    from pde_module_amd.FDM.boundary.tables import CHUNK, chunk_ops, chunk_transpose, transpose_table
    name, u, bc = _chunk_cases()[case]
    ops = bc.resolved_ops()
    ch = chunk_ops(ops)
    assert ch is not None
    cops, starts = ch
    assert sorted(map(tuple, cops.tolist())) == sorted(map(tuple, ops.tolist()))
    assert np.all(np.diff(starts) <= CHUNK) and starts[-1] == len(ops)
    x = np.random.default_rng(1).uniform(-1, 1, u.shape).astype(np.float32)
    flat = x.ravel().copy()
    for c in range(len(starts) - 1):
        part = cops[starts[c]:starts[c + 1]]
        vals = apply_ops_host(flat, part).ravel()[part["dst"]]   # reads of the current field
        flat[part["dst"]] = vals
    assert same_bits(flat.reshape(x.shape), apply_ops_host(x, ops))

    T = transpose_table(ops)
    C = chunk_transpose(T)
    assert C is not None and np.all(np.diff(C["item_start"]) <= CHUNK)
    gflat = np.random.default_rng(2).uniform(-1, 1, x.size).astype(np.float32)
    want = O.boundary_linear_adjoint(gflat.copy(), [(int(o["dst"]), int(o["src"]),
                                      {0: 0, 1: 1, 2: -1, 3: 1}[int(o["kind"])]) for o in ops])
    g = gflat.copy()
    for c in range(len(C["item_start"]) - 1):
        i0, i1 = C["item_start"][c], C["item_start"][c + 1]
        staged = []
        for i in range(i0, i1):
            acc = np.float32(0.0) if C["flags"][i] & 1 else g[C["src_idx"][i]]
            for q in range(C["off"][i], C["off"][i + 1]):
                acc = np.float32(acc + C["csign"][q] * g[C["cdst"][q]])
            staged.append(acc)
        g[C["src_idx"][i0:i1]] = staged
        g[C["zero_idx"][C["zero_start"][c]:C["zero_start"][c + 1]]] = 0.0
    assert same_bits(g, want.ravel())


def test_chunking_refuses_oversized_components():
    from pde_module_amd._native import BC_COPY, BC_OP_DTYPE
    from pde_module_amd.FDM.boundary.tables import chunk_ops
    ops = np.zeros(300, BC_OP_DTYPE)
    ops["dst"] = np.arange(300)
    ops["src"] = np.arange(1, 301)          # a chain: each op reads the next one's target
    ops["kind"] = BC_COPY
    assert chunk_ops(ops) is None
    assert chunk_ops(ops[:200]) is not None


def test_caller_time_array_is_checked_before_any_launch():
    """The one-launch steps read a caller's time array on the device: a
    wrong dtype, size or device is refused on the host (the same check as the
    module call's), never handed to a kernel."""
    u = field((6, 6, 1), 1)
    bc = GridBoundary(u, 0.1, 1)
    bc.dirichlet_BC("ALL", get_ramp_func(1.0, 2.0))
    for bad in (torch.zeros(1, dtype=torch.float64), torch.zeros(2),
                torch.zeros(1, device="meta")):
        with pytest.raises(AssertionError):
            bc._time_tensor(u, bad)
    good = torch.zeros(1)
    assert bc._time_tensor(u, good) is good


def test_boundary_hooks_are_registered():
    """The time and params set-up of both boundary classes runs as a
    before-forward hook (a decorator left on another method would make every
    ramp read a stale time)."""
    for cls in (GridBoundary, LegacyGridBoundary):
        fn = getattr(cls, "set_default_params")
        assert hasattr(fn, "_before_forward_order"), cls
    assert not hasattr(GridBoundary.__dict__["_check_time_array"], "_before_forward_order")
