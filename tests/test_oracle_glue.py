"""Pins of the node-wise glue restatements (outer, scalar x vector, matrix x
vector, vector / scalar, linear combinations, shifted scale) against what the
arithmetic fixes: products of two FP32 numbers and quotients are exact or
correctly rounded in FP64, so one FP64 operation rounded to FP32 is the FP32
result; integer-valued inputs make every sum exact; a non-symmetric matrix
catches a transposed operand."""
import numpy as np

from oracle import np_oracle as O

F32 = np.float32


def _r(shape, seed):
    return np.random.default_rng(seed).uniform(-1, 1, shape).astype(F32)


def test_outer_exact_products():
    a, b = _r((5, 4, 1, 2), 0), _r((5, 4, 1, 3), 1)
    got = O.outer(a, b)
    assert got.shape == (5, 4, 1, 2, 3) and got.dtype == F32
    want = np.einsum("...i,...j->...ij", a.astype(np.float64), b.astype(np.float64)).astype(F32)
    assert np.array_equal(got, want)
    assert got[2, 1, 0, 1, 2] == a[2, 1, 0, 1] * b[2, 1, 0, 2]


def test_scalar_multiply_and_divide_correctly_rounded():
    s = _r((6, 5, 1, 1), 2) + F32(2.0)
    v = _r((6, 5, 1, 2), 3)
    m = _r((6, 5, 1, 2, 2), 4)
    assert np.array_equal(O.scalar_vector_multiply(s, v),
                          (s.astype(np.float64) * v.astype(np.float64)).astype(F32))
    assert np.array_equal(O.scalar_vector_multiply(s, m),
                          (s[..., None].astype(np.float64) * m).astype(F32))
    assert np.array_equal(O.divide_by_scalar(v, s),
                          (v.astype(np.float64) / s.astype(np.float64)).astype(F32))


def test_matvec_integer_exact_and_orientation():
    rng = np.random.default_rng(5)
    m = rng.integers(-8, 9, (4, 3, 1, 2, 3)).astype(F32)
    v = rng.integers(-8, 9, (4, 3, 1, 3)).astype(F32)
    got = O.matvec(m, v)
    assert got.shape == (4, 3, 1, 2)
    assert np.array_equal(got, np.einsum("...ij,...j->...i", m, v))
    # one node by hand: row i uses m[i][:], not the column
    M = np.array([[1, 2], [3, 4]], F32).reshape(1, 1, 1, 2, 2)
    x = np.array([10, 100], F32).reshape(1, 1, 1, 2)
    assert O.matvec(M, x).ravel().tolist() == [210.0, 430.0]


def test_matvec_fp32_within_bound_of_fp64():
    m, v = _r((7, 6, 1, 3, 3), 6), _r((7, 6, 1, 3), 7)
    got = O.matvec(m, v).astype(np.float64)
    ref = np.einsum("...ij,...j->...i", m.astype(np.float64), v.astype(np.float64))
    bound = 3 * np.finfo(F32).eps * np.einsum("...ij,...j->...i", np.abs(m), np.abs(v))
    assert np.all(np.abs(got - ref) <= bound)


def test_linear_combination_is_the_operator_chain():
    a, b, c, d = (_r((9, 8, 1, 2), s) for s in range(10, 14))
    got = O.linear_combination((1, a), (-1, b), (1, c), (1, d))
    assert np.array_equal(got, ((a - b) + c) + d)
    ai = np.random.default_rng(1).integers(-50, 50, (3, 3, 1, 1)).astype(F32)
    bi = np.random.default_rng(2).integers(-50, 50, (3, 3, 1, 1)).astype(F32)
    assert np.array_equal(O.linear_combination((2, ai), (-3, bi)), 2 * ai - 3 * bi)


def test_shifted_scale_components():
    x = _r((4, 4, 1, 3), 20)
    y = O.shifted_scale(x, 100.0, 1.0)
    assert np.array_equal(y, F32(100.0) * (x - F32(1.0)))
    z = O.shifted_scale(x, 100.0, 1.0, comp=1)
    assert np.array_equal(z[..., 0], x[..., 0]) and np.array_equal(z[..., 2], x[..., 2])
    assert np.array_equal(z[..., 1], y[..., 1])
