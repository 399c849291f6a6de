"""The allocation-free in-place boundary (pde_bc_apply_inplace_noalloc_f32),
called through the C ABI, against pde_bc_apply_inplace_f32 with a caller's
staging array: identical bits on two calls in a row, the second with a
smaller table, both above the 4096 ops that one workgroup stages in LDS (so
the file-scope scratch array is used), then a table small enough for the LDS
path."""
import numpy as np
import pytest
import torch

import pde_module_amd._native as N
from pde_module_amd.deprecated import GridBoundary as LegacyGridBoundary

pytestmark = pytest.mark.gpu


def _case(dev, shape, seed):
    rng = np.random.default_rng(seed)
    u = torch.from_numpy(rng.uniform(-1, 1, shape + (1,)).astype(np.float32)).to(dev)
    bc = LegacyGridBoundary(u, 0.05, 1)
    bc.vonNeumann_BC("-X", 0.5)
    bc.dirichlet_BC("+Y", 0.25)
    ops = np.ascontiguousarray(bc.resolved_ops())   # unchunked order: one op per dst
    return u, ops


def test_no_alloc_boundary_matches_staged_boundary(dev):
    for shape, seed in (((70, 66, 64), 1), ((40, 36, 30), 2), ((30, 30, 1), 3)):
        print(f"allocation-free boundary {shape}", flush=True)
        u, ops = _case(dev, shape, seed)
        nops = len(ops)
        if shape[2] > 1:
            assert nops > 4096          # the scratch path, not the LDS one
        ops_dev = torch.from_numpy(ops.view(np.uint8).copy()).to(dev)
        a, b = u.clone(), u.clone()
        staging = torch.empty(max(nops, 1), dtype=torch.float32, device=dev)
        st = N.stream_of(a)
        N.call("pde_bc_apply_inplace_f32", a.data_ptr(), ops_dev.data_ptr(), nops, None,
               staging.data_ptr(), st)
        N.call("pde_bc_apply_inplace_noalloc_f32", b.data_ptr(), ops_dev.data_ptr(), nops,
               None, st)
        torch.cuda.synchronize()
        assert torch.equal(a.view(torch.int32), b.view(torch.int32))
        assert not torch.equal(a, u)    # the boundary did write
