"""The C-ABI library loads and exports every entry point include/pde_hip.h
declares; the op-record layout matches the header (CPU only: no compute)."""
import ctypes
import glob
import os
import re
import struct
import subprocess

import numpy as np
import pytest

import pde_module_amd._native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))
           if f.endswith(".h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?(?:int|char\s*\*|const char\s*\*|void)\s*\*?\s*(pde_\w+)\s*\(",
                             text, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_the_bound_entry_points():
    names = declared_functions()
    assert names, "no declarations found"
    assert names == set(N.SIGNATURES), names ^ set(N.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = N.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", N.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert declared_functions() <= exported


def offload_bundle_targets(path):
    """Target triples of every clang offload bundle embedded in a shared
    library (one per HIP translation unit): the bundle header is the magic
    string, a little-endian u64 entry count, then per entry u64 offset, u64
    size, u64 triple length and the triple."""
    data = open(path, "rb").read()
    magic = b"__CLANG_OFFLOAD_BUNDLE__"
    triples, nbundles = set(), 0
    for m in re.finditer(re.escape(magic), data):
        nbundles += 1
        q = m.end()
        (n,) = struct.unpack_from("<Q", data, q)
        q += 8
        for _ in range(n):
            _, _, tl = struct.unpack_from("<QQQ", data, q)
            q += 24
            triples.add(data[q:q + tl].decode())
            q += tl
    return nbundles, triples


def test_library_has_gfx950_code_only():
    nbundles, triples = offload_bundle_targets(N.LIB_PATH)
    nsources = len(glob.glob(os.path.join(ROOT, "pde_module_amd", "csrc", "*.hip")))
    assert nbundles == nsources, (nbundles, nsources)
    device = {t for t in triples if "amdgcn" in t}
    assert device == {"hipv4-amdgcn-amd-amdhsa--gfx950"}, triples
    assert triples - device == {"host-x86_64-unknown-linux-gnu-"}, triples


def test_version_and_status_strings():
    lib = N.load()
    assert lib.pde_version() >= 100
    assert lib.pde_status_string(0) == b"ok"
    assert b"invalid" in lib.pde_status_string(-1)


def test_bc_op_record_layout():
    text = open(os.path.join(ROOT, "include", "pde_hip.h")).read()
    body = re.search(r"typedef struct pde_bc_op \{(.*?)\} pde_bc_op;", text, flags=re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    union = re.search(r"union \{(.*?)\};", body, flags=re.S).group(1)
    members = re.findall(r"(int32_t|float)\s+(\w+);", union)
    assert [f for _, f in members] == ["value", "slot"]      # 4 bytes shared
    flat = re.sub(r"union \{.*?\};", "float value;", body, flags=re.S)
    fields = re.findall(r"(int64_t|int32_t|int16_t|float)\s+(\w+);", flat)
    assert [f for _, f in fields] == list(N.BC_OP_DTYPE.names)
    sizes = {"int64_t": 8, "int32_t": 4, "int16_t": 2, "float": 4}
    assert sum(sizes[t] for t, _ in fields) == N.BC_OP_DTYPE.itemsize == 32
    kinds = dict(re.findall(r"#define (PDE_BCV?_\w+) (\d+)", text))
    assert int(kinds["PDE_BC_SET"]) == N.BC_SET and int(kinds["PDE_BC_GHOST_N"]) == N.BC_GHOST_N
    assert (int(kinds["PDE_BCV_CONST"]), int(kinds["PDE_BCV_SINE"]),
            int(kinds["PDE_BCV_ARRAY"])) == (N.BCV_CONST, N.BCV_SINE, N.BCV_ARRAY)
    assert np.array_equal(N.bc_slots(np.array([(0, 0, 0, 2, N.bc_slot_bits([7])[0], 0, 0)],
                                              dtype=N.BC_OP_DTYPE)), [7])


def test_multisweep_jacobi_arguments_need_no_gpu():
    """The multi-sweep Jacobi entry points check the sweep count (1 ..
    PDE_JACOBI_MAX_SWEEPS), aliasing and the 2D shape before any launch."""
    text = open(os.path.join(ROOT, "include", "pde_hip.h")).read()
    assert int(re.search(r"#define PDE_JACOBI_MAX_SWEEPS (\d+)", text).group(1)) == \
        N.JACOBI_MAX_SWEEPS
    lib = N.load()
    a, b, c = (ctypes.c_void_p(16 * i) for i in (1, 2, 3))   # never dereferenced
    fwd = lib.pde_jacobi_neumann0_multisweep_f32
    adj = lib.pde_jacobi_neumann0_adjoint_multisweep_f32
    assert adj(a, None, b, c, 1, 66, 66, 1, 1, 1.0, 0.25, 4, 2, None) == -1   # accumulate 0/1
    assert adj(a, b, b, c, 1, 66, 66, 1, 1, 1.0, 0.25, 4, 1, None) == -1      # g2 aliases gp
    for fn in (fwd, lambda x0, *x: adj(x0, None, *x[:-1], 1, x[-1])):
        assert fn(a, b, c, 1, 66, 66, 1, 1, 1.0, 0.25, 0, None) == -1
        assert fn(a, b, c, 1, 66, 66, 1, 1, 1.0, 0.25, N.JACOBI_MAX_SWEEPS + 1, None) == -1
        assert fn(a, b, a, 1, 66, 66, 1, 1, 1.0, 0.25, 4, None) == -1     # out aliases in
        assert fn(a, b, c, 0, 66, 66, 1, 1, 1.0, 0.25, 4, None) == -1
        assert fn(a, b, c, 1, 66, 66, 66, 1, 1.0, 0.25, 4, None) == -2     # 3D
        assert fn(a, b, c, 1, 66, 3, 1, 1, 1.0, 0.25, 4, None) == -2       # 1 interior column
        assert fn(a, b, c, 70000, 66, 66, 1, 1, 1.0, 0.25, 4, None) == -2


def test_batched_cavity_arguments_need_no_gpu():
    """Batch counts outside 1..65535 and 3D grids are refused before a launch."""
    lib = N.load()
    a, b, c = (ctypes.c_void_p(16 * i) for i in (1, 2, 3))   # never dereferenced
    f = lib.pde_divergence_batch_f32
    assert f(a, b, 0, 1, 40, 40, 1, 1, 1.0, 0.0, 1.0, 1.0, None) == -1
    assert f(a, b, 70000, 1, 40, 40, 1, 1, 1.0, 0.0, 1.0, 1.0, None) == -2
    assert f(a, b, 2, 40, 40, 40, 1, 1, 1.0, 0.0, 1.0, 1.0, None) == -2
    assert lib.pde_cavity_correct_batch_f32(a, b, c, 0, 40, 40, 1, 1, 1.0, 0.0, 1.0, 0.1,
                                            None) == -1
    assert lib.pde_cavity_predict_batch_f32(a, b, 70000, 40, 40, 1, 1, 1.0, -2.0, 1.0, 1.0,
                                            0.0, 1.0, 0.01, 0.1, None) == -2
    assert lib.pde_cavity_adjoint_div_batch_f32(a, b, a, 2, 40, 40, 1, 1, 1.0, 0.0, 1.0, 1.0,
                                                None) == -1       # gus2 aliases grhs
    assert lib.pde_cavity_adjoint_correct_batch_f32(a, a, 2, 40, 40, 1, 1, 1.0, 0.0, 1.0, 0.1,
                                                    None) == -1   # gpf aliases gu_out
    assert lib.pde_cavity_adjoint_correct_f32(a, a, 40, 40, 1, 1, 1.0, 0.0, 1.0, 0.1,
                                              None) == -1


def test_argument_errors_need_no_gpu():
    """Entry points validate arguments before touching the device."""
    lib = N.load()
    assert lib.pde_laplacian_f32(None, None, 8, 8, 8, 1, 1, 1.0, -2.0, 1.0, 1.0, None) == -1
    fake = ctypes.c_void_p(16)  # never dereferenced: the shape check fails first
    assert lib.pde_laplacian_f32(fake, fake, 2, 8, 8, 1, 1, 1.0, -2.0, 1.0, 1.0, None) == -2
    assert lib.pde_bc_apply_f32(fake, fake, 10, None, 5, None, None) == -1
    assert lib.pde_wave_step_f32(fake, fake, None, fake, fake, 8, 8, 1, 1, 1.0, -2.0, 1.0,
                                 1.0, 0.1, 0, None) == -1


def test_ring_plan_record_layout_and_op_count_check():
    """The plan record matches the header's struct, and both ring entry points
    refuse a plan built for another op table (its nops differs from the table
    passed) on the host, before any launch."""
    text = open(os.path.join(ROOT, "include", "pde_hip.h")).read()
    body = re.search(r"typedef struct pde_ring_plan \{(.*?)\} pde_ring_plan;", text,
                     flags=re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = re.findall(r"(\w+)\s*[,;]", body)
    assert names == list(N.RING_PLAN_DTYPE.names)
    import torch
    from pde_module_amd.deprecated import GridBoundary as LegacyGridBoundary
    from pde_module_amd.FDM.boundary.tables import chunk_ops
    from pde_module_amd.ring_plan import build_ring_plan
    ops = chunk_ops(LegacyGridBoundary(torch.zeros((34, 34, 1, 1)), 0.01, 1).resolved_ops())[0]
    plan = build_ring_plan(ops, 34, 34, 1)
    assert plan["nops"] == len(ops)
    lib = N.load()
    pieces = np.zeros(4, np.uint8)
    rec = np.zeros(1, N.RING_PLAN_DTYPE)
    fake = pieces.ctypes.data          # never dereferenced: the op count check fails first
    rec[0] = (plan["depth"], len(plan["pieces"]), len(ops) + 1, fake, fake, fake, fake, fake,
              fake)
    u, un, ops_p, t, save = (ctypes.c_void_p(16 * k) for k in (1, 2, 3, 4, 5))
    assert lib.pde_heat_step_ring_f32(u, un, 34, 34, 1, 1, ops_p, len(ops), None, 1.0, -2.0,
                                      1.0, 0.01, 0.1, rec.ctypes.data, None) == -1
    step = np.zeros(2, np.int64)
    assert lib.pde_wave_reverse_adjoint_pair_f32(
        u, un, ctypes.c_void_p(48), ctypes.c_void_p(64), None, ctypes.c_void_p(80),
        ctypes.c_void_p(96), ctypes.c_void_p(112), ctypes.c_void_p(128), None, 34, 34, 1, 1,
        1.0, -2.0, 1.0, 1.0, 0.1, step.ctypes.data, step[1:].ctypes.data, None, None, save,
        len(ops), rec.ctypes.data, None) == -1


def test_allocation_free_boundary_refuses_oversize_tables_on_the_host():
    """pde_bc_apply_inplace_noalloc_f32 returns PDE_ETOOBIG (-3) for a table
    larger than its fixed scratch array before touching the device."""
    lib = N.load()
    fake = ctypes.c_void_p(16)
    assert lib.pde_bc_apply_inplace_noalloc_f32(fake, fake, (1 << 20) + 1, None, None) == -3
    assert b"scratch" in lib.pde_status_string(-3)
    assert lib.pde_bc_apply_inplace_noalloc_f32(None, fake, 10, None, None) == -1
    assert lib.pde_bc_apply_inplace_noalloc_f32(fake, None, 10, None, None) == -1
