"""The two FP32 restatements (C and NumPy) agree bit for bit, and both stay
within FP32 rounding of the FP64 oracle (CPU only)."""
import numpy as np
import pytest

from conftest import same_bits
from oracle import np_oracle as O

SHAPES = [(10, 12, 1), (6, 7, 8), (9, 1, 1), (1, 6, 5), (5, 1, 7)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("C", [1, 3])
@pytest.mark.parametrize("g", [1, 2])
def test_operators_c_equals_numpy_fp32(corc, shape, C, g):
    rng = np.random.default_rng(hash((shape, C, g)) % 2**32)
    u = rng.uniform(-1, 1, shape + (C,)).astype(np.float32)
    dx = 0.1 + 0.01 * C
    wl, wc = O.laplacian_weights(dx), O.central_weights(dx)
    assert same_bits(O.laplacian(u, wl, 0.7, g), corc.laplacian(u, wl, np.float32(0.7), g))
    assert same_bits(O.grad(u, wc, 1.3, g), corc.grad(u, wc, np.float32(1.3), g))
    D = len(O.eligible_axes(shape))
    v = rng.uniform(-1, 1, shape + (D,)).astype(np.float32)
    assert same_bits(O.divergence(v, wc, -0.5, g), corc.divergence(v, wc, np.float32(-0.5), g))
    m = rng.uniform(-1, 1, shape + (C, D)).astype(np.float32)
    assert same_bits(O.divergence(m, wc, 2.0, g, True), corc.divergence(m, wc, np.float32(2.0), g, True))


def test_openmp_threads_do_not_change_results(corc):
    rng = np.random.default_rng(0)
    u = rng.uniform(-1, 1, (34, 33, 35, 1)).astype(np.float32)
    w = O.laplacian_weights(0.05)
    a = corc.laplacian(u, w, np.float32(1.0), 1, nthreads=1)
    b = corc.laplacian(u, w, np.float32(1.0), 1, nthreads=4)
    assert same_bits(a, b)


@pytest.mark.parametrize("shape", [(12, 11, 1), (7, 8, 9)])
def test_fp32_within_rounding_of_fp64(corc, shape):
    rng = np.random.default_rng(1)
    u64 = rng.uniform(-1, 1, shape + (1,))
    u32 = u64.astype(np.float32)
    dx = 1.0 / 37
    lap32 = corc.laplacian(u32, O.laplacian_weights(dx), np.float32(1.0), 1)
    lap64 = O.laplacian(u32.astype(np.float64), O.laplacian_weights(dx, np.float64), 1.0, 1)
    inner = O.interior(shape, 1)
    scale = np.abs(O.laplacian_weights(dx, np.float64)[1]) * O.eligible_axes(shape).__len__() * 2
    err = np.max(np.abs(lap32[inner] - lap64[inner])) / scale
    assert err < 1e-6


def test_boundaries_c_equals_numpy(corc):
    rng = np.random.default_rng(2)
    for shape in [(9, 8, 1), (6, 7, 5)]:
        for C in (1, 2, 3):
            u = rng.uniform(-1, 1, shape + (C,)).astype(np.float32)
            ijk, groups, inter = O.legacy_boundary_layout(shape, 1)
            bt = rng.integers(0, 3, (len(ijk), C)).astype(np.int8)
            bv = rng.uniform(-2, 2, (len(ijk), C)).astype(np.float32)
            a = O.bc_legacy(u, ijk, inter, bt, bv, np.float32(0.1))
            b = corc.bc_legacy(u, ijk, inter, bt, bv, 0.1)
            assert same_bits(a, b)
            fg = O.functional_groups(shape, 1)
            grps = [dict(ids=fg["ALL"], type=1, comps=list(range(C)), const=0.0),
                    dict(ids=fg["+Y"], type=1, comps=[0], const=1.0, tmax=0.3),
                    dict(ids=fg["-X"], type=2, comps=[C - 1], const=-0.5)]
            for t in (0.0, 0.1, 1.0):
                a = O.bc_functional(u, 1, grps, np.float32(0.1), t=t)
                b = corc.bc_functional(u, 1, grps, 0.1, t=t)
                assert same_bits(a, b)


def test_periodic_fill_c_equals_numpy(corc):
    rng = np.random.default_rng(3)
    for shape, g in [((8, 9, 10), 1), ((9, 10, 1), 2)]:
        u = rng.uniform(-1, 1, shape + (2,)).astype(np.float32)
        a = O.periodic_fill(u.copy(), g)
        b = corc.periodic_fill(u.copy(), g)
        assert same_bits(a, b)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("g", [1, 2])
def test_wave_step_c_equals_numpy(corc, shape, g):
    """The C wave update (the bench's CPU wave step) against the NumPy one, bit
    for bit: per-node c2 and none, first step and later steps, 1 and all
    threads."""
    rng = np.random.default_rng(hash((shape, g, "wave")) % 2**32)
    u2 = rng.uniform(-1, 1, shape + (1,)).astype(np.float32)
    v = rng.uniform(-1, 1, shape + (1,)).astype(np.float32)
    c2 = rng.uniform(0.5, 1.5, shape + (1,)).astype(np.float32)
    w = O.laplacian_weights(0.07)
    for cc in (c2, None):
        for first in (True, False):
            want = O.wave_step_core(u2, v, cc, w, np.float32(0.9), g, np.float32(0.01), first)
            for threads in (1, corc.max_threads()):
                got = corc.wave_step_core(u2, v, cc, w, 0.9, g, 0.01, first, nthreads=threads)
                assert same_bits(got[0], want[0]) and same_bits(got[1], want[1])
