"""Writes tests/golden/fullsize_counts.json: the oracle's (the CPU
restatement's) iteration counts and final residual estimates for the
benchmark configurations at their full sizes, so the GPU tests can compare
the device solvers with them without re-running minute-long CPU solves.

    python tests/golden/make_fullsize_counts.py [config1 config2a ...]

(all four, about 5 minutes, when no name is given; named ones are
recomputed and merged into the existing file).

Inputs are the benchmark's: RHS splitmix64 uniform[-1,1] (seed 42), x0 = 0.
  config 1   poisson_2d(256) (Neumann, matgen.hpp:173-215), mean-free RHS,
             Jacobi-PCG to 1e-8, restart_iterations = max = 20000
             (conjugate_gradient.hpp:122-312)
  config 2a  7-point Poisson 192^3, Jacobi-PCG to 1e-8 (same solver)
  config 3   3D upwind convection-diffusion 128^3 (Peclet from seed 42),
             FP64 GMRES(50) with the exact ILU(0) apply as the right
             preconditioner (the residual GMRES monitors is then b - A x)
             to 1e-10 (gmres.hpp:99-375, ilu.hpp:35-179)
  config 5   2D anisotropic Q1 diffusion 1024^2 (seed 4242, sigma 1.5),
             Jacobi-PCG to 1e-8
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402


def _mix(seed, k):
    m = (1 << 64) - 1
    z = (seed + (k + 1) * 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31)


def convdiff_peclet(seed=42):
    """Peclet U[10, 100] drawn from splitmix64(seed ^ 0x5045), element 0."""
    return 10.0 + 90.0 * ((_mix(seed ^ 0x5045, 0) >> 11) / 2.0 ** 53)


def _entry(s, x, A, b, secs):
    return {"iterations": s.computed_iterations, "converged": s.converged,
            "residual_norm": s.residual_norm,
            "true_relres": O.true_relres(A, b, x), "oracle_s": round(secs, 1)}


def config1():
    A = O.poisson_2d(256, 256, 0.0)
    b = O.uniform_pm1(42, A.nrows)
    b = b - b.mean()
    t = time.perf_counter()
    x, s = O.cg(A, b, tol=1e-8, restart_iterations=20000,
                max_iterations=20000, left="jacobi")
    return "config1_poisson2d_256_jacobi_pcg", _entry(
        s, x, A, b, time.perf_counter() - t)


def config2a():
    A = O.poisson_3d(192)
    b = O.uniform_pm1(42, A.nrows)
    t = time.perf_counter()
    x, s = O.cg(A, b, tol=1e-8, restart_iterations=20000,
                max_iterations=20000, left="jacobi")
    return "config2a_poisson3d_192_jacobi_pcg", _entry(
        s, x, A, b, time.perf_counter() - t)


def config3():
    pe = convdiff_peclet(42)
    h = 1.0 / 129
    v = np.array([1.0, 0.5, 0.25])
    c = pe * h * v / np.linalg.norm(v)
    A = O.convdiff_3d(128, 128, 128, 1.0, *c)
    b = O.uniform_pm1(42, A.nrows)
    t = time.perf_counter()
    x, s = O.gmres(A, b, tol=1e-10, restart_dimension=50, max_iterations=5000,
                   right="ilu")
    e = _entry(s, x, A, b, time.perf_counter() - t)
    e["peclet"] = pe
    return "config3_convdiff_128_gmres50_right_ilu0_fp64", e


def config5():
    A = O.aniso_q1_2d(1024, 4242, 1.5)
    b = O.uniform_pm1(42, A.nrows)
    t = time.perf_counter()
    x, s = O.cg(A, b, tol=1e-8, restart_iterations=50000,
                max_iterations=50000, left="jacobi")
    return "config5_aniso_q1_1024_jacobi_pcg", _entry(
        s, x, A, b, time.perf_counter() - t)


RUNS = {"config1": config1, "config2a": config2a, "config3": config3,
        "config5": config5}


def main():
    path = os.path.join(HERE, "fullsize_counts.json")
    names = sys.argv[1:] or list(RUNS)
    out = {}
    if sys.argv[1:] and os.path.exists(path):
        with open(path) as f:
            out = json.load(f)
    for nm in names:
        key, e = RUNS[nm]()
        out[key] = e
        print(key, e, flush=True)
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()
