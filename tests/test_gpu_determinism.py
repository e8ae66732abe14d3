"""Every reduction on the device sums its partials in one fixed order --
through a one-block finisher or in the prologue of the next kernel, where
every block derives the same scalars -- so a repeated solve returns the same
bits: iterates, iteration counts and residual histories, for matrices on the
one-kernel hand-off path (few SpMV row blocks) and on the finisher path."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _twice(fn):
    a, b = fn(), fn()
    return a, b


@pytest.mark.parametrize("size", ["handoff", "finisher"])
@pytest.mark.parametrize("method", ["cg", "bicgstab", "cgs", "bicg"])
def test_repeated_solves_are_bit_identical(O, spk, method, size):
    if method == "cg":
        A = spk.poisson_3d(40, 40, 40) if size == "handoff" else spk.poisson_3d(136, 136, 136)
    else:
        A = (spk.convdiff_3d(40, 40, 40, 1.0, 0.3, 0.2, 0.1) if size == "handoff"
             else spk.convdiff_3d(136, 136, 136, 1.0, 0.3, 0.2, 0.1))
    b = spk.uniform_pm1(7, A.nrows)
    J = spk.jacobi(A)

    def run():
        x = spk.DeviceVector.zeros(A.nrows)
        if method == "cg":
            s = spk.conjugate_gradient(b, x, spk.CGConfig(1e-9, 5000, 120, True), A, J)
        elif method == "bicgstab":
            s = spk.bicgstab(b, x, spk.BicgstabConfig(1e-9, 120, True), A, J)
        elif method == "cgs":
            s = spk.cgs(b, x, spk.CgsConfig(1e-9, 120, True), A, J)
        else:
            s = spk.bicg(b, x, spk.BicgConfig(1e-9, 120, True), A, A.transpose(), J)
        return x.numpy(), s.computed_iterations, np.array(s.iteration_residuals)

    (x1, k1, h1), (x2, k2, h2) = _twice(run)
    assert k1 == k2
    assert np.array_equal(x1.view(np.uint64), x2.view(np.uint64))
    assert np.array_equal(h1.view(np.uint64), h2.view(np.uint64))


def test_batched_repeat_is_bit_identical(spk):
    B = spk.BatchedSystems.diffusion(64, 3, 16, 96)
    b = spk.uniform_pm1(9, B.total_rows)
    cfg = spk.BatchConfig(1e-8, 1e-2, 30, 5000, spk.F32)
    xs = []
    for _ in range(2):
        x = spk.DeviceVector.zeros(B.total_rows)
        B.solve(cfg, b, x)
        xs.append((x.numpy(), B.results()[0].copy()))
    assert np.array_equal(xs[0][0].view(np.uint64), xs[1][0].view(np.uint64))
    assert np.array_equal(xs[0][1], xs[1][1])
