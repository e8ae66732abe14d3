"""spk_dot / spk_axpby (the BLAS-1 primitives behind the solvers' inner
products, std::inner_product at conjugate_gradient.hpp:194,236) against
numpy at the 192^3 Poisson vector length, FP64 and FP32.

The device dot sums in a fixed tree order, not left to right, so it is
compared with the exactly rounded sum (math.fsum of the exact products)
within a rounding bound; axpby may contract a*x + b*y into one fused
multiply-add, so it is compared within two units in the last place of the
terms."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 192 ** 3  # 7,077,888: one vector of the 192^3 Poisson solve


def _vectors(dtype, seed=11):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1.0, 1.0, N).astype(dtype)
    y = rng.uniform(-1.0, 1.0, N).astype(dtype)
    return x, y


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_dot_matches_exact_sum(spk, dtype):
    x, y = _vectors(dtype)
    dx = spk.DeviceVector.from_numpy(x)
    dy = spk.DeviceVector.from_numpy(y)
    got = spk.dot(dx, dy)
    xd, yd = x.astype(np.float64), y.astype(np.float64)
    prods = xd * yd  # exact for float32 inputs, within eps/2 for float64
    exact = math.fsum(prods)
    scale = float(np.abs(prods).sum())
    # each thread sums ~14 strided terms, then a ~20-level tree: the FP64
    # error stays below ~40 eps * sum|x y|; 128 eps leaves room, while a
    # dropped or doubled term (|x_i y_i| ~ 0.25 against 5e-11 here) fails
    tol = 128 * np.finfo(np.float64).eps * scale
    assert abs(got - exact) <= tol, (got, exact, tol)
    assert got == spk.dot(dx, dy)  # fixed order: repeatable bit for bit


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_axpby_matches_numpy(spk, dtype):
    x, y = _vectors(dtype, seed=12)
    a, b = 0.75, -1.25
    dx = spk.DeviceVector.from_numpy(x)
    dy = spk.DeviceVector.from_numpy(y)
    spk.axpby(a, dx, b, dy)
    got = dy.numpy()
    ta, tb = dtype(a) * x, dtype(b) * y
    ref = ta + tb
    eps = np.finfo(dtype).eps
    tol = 2 * eps * (np.abs(ta) + np.abs(tb))
    assert np.all(np.abs(got.astype(np.float64) - ref.astype(np.float64))
                  <= tol.astype(np.float64)), "axpby differs beyond rounding"
    # x is untouched and the zero-length call is a no-op
    assert np.array_equal(dx.numpy(), x)
    e = spk.DeviceVector(0, dtype)
    spk.axpby(1.0, e, 1.0, e)
    assert spk.dot(e, e) == 0.0


def test_dot_rejects_mismatched_vectors(spk):
    a = spk.DeviceVector.zeros(8)
    b = spk.DeviceVector.zeros(8, np.float32)
    with pytest.raises(spk.InvalidArgument):
        spk.dot(a, b)
    with pytest.raises(spk.InvalidArgument):
        spk.axpby(1.0, a, 1.0, spk.DeviceVector.zeros(9))
