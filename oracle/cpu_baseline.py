"""CPU baseline legs: the oracle (a CPU restatement of sparkit's solve path;
test infrastructure) timed on one host core, for the benchmark's report.

    python -m oracle.cpu_baseline LEG[,LEG...] --core K

pins itself to core K (what `taskset -c K` does) before it touches any data,
runs the legs one after another and prints one JSON line per leg.  The
operator slots allocate and copy per call as the reference's multiply /
ilu_apply do (sparse_blas.hpp:25, ilu.hpp:173-179), so the timing follows
the reference's memory traffic.
The library is compiled with the reference's Release flags (-O3 -DNDEBUG,
CMakePresets.json:15) and no FMA contraction.

Legs (inputs are the benchmark's; RHS splitmix64 uniform[-1,1], x0 = 0):
  p2d        poisson_2d(256), mean-free RHS, Jacobi-PCG to 1e-8, full solve
  poisson192 7-point Poisson 192^3, Jacobi-PCG to 1e-8, full solve
  powerlaw   power-law SPD: n = 262,144 sample of the same law (the 4.19M
             matrix takes minutes to build on one core), 11 Jacobi-PCG
             iterations (the full matrix's count), scaled to the full
             matrix by nnz -- labelled extrapolated
  convdiff   3D upwind convection-diffusion 128^3, FP64 GMRES(50) with the
             exact ILU(0) apply in the right slot to 1e-10, full solve
             (factorization timed apart)
  batched    64 of the 4096 batched diffusion systems (seeded choice),
             Jacobi-PCG to 1e-8 each, scaled to the batch by nnz --
             labelled extrapolated
  aniso      2D anisotropic Q1 diffusion 1024^2, Jacobi-PCG to 1e-8, full
             solve
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from oracle import oracle as O  # noqa: E402


def b_spmv(n, nnz, s=8):
    """Algorithmic bytes of one SpMV, int32 indices (the GPU byte model)."""
    return nnz * (s + 4) + 4 * (n + 1) + 2 * s * n


def b_pcg(n, nnz, s=8):
    return b_spmv(n, nnz, s) + 11 * s * n


def _mix(seed, k):
    m = (1 << 64) - 1
    z = (seed + (k + 1) * 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31)


def _pcg(A, b, tol, max_it):
    t0 = time.perf_counter()
    x, s = O.cg(A, b, tol=tol, restart_iterations=max_it,
                max_iterations=max_it, left="jacobi", ref_calls=True)
    return x, s, time.perf_counter() - t0


def leg_p2d():
    A = O.poisson_2d(256, 256, 0.0)
    b = O.uniform_pm1(42, A.nrows)
    b = b - b.mean()
    x, s, dt = _pcg(A, b, 1e-8, 20000)
    return {"input": "poisson_2d(256), mean-free RHS", "solver":
            "Jacobi-PCG FP64 to 1e-8", "full_solve": True, "s": round(dt, 3),
            "iterations": s.computed_iterations, "converged": s.converged,
            "true_relres": O.true_relres(A, b, x)}


def leg_poisson192():
    A = O.poisson_3d(192)
    b = O.uniform_pm1(42, A.nrows)
    x, s, dt = _pcg(A, b, 1e-8, 20000)
    it = s.computed_iterations
    return {"input": "3D 7-point Poisson 192^3", "solver":
            "Jacobi-PCG FP64 to 1e-8", "full_solve": True, "s": round(dt, 3),
            "iterations": it, "converged": s.converged,
            "true_relres": O.true_relres(A, b, x),
            "rows": A.nrows, "nnz": A.nnz,
            "GBs": round(it * b_pcg(A.nrows, A.nnz) / dt / 1e9, 3)}


# power-law matrix at the benchmark size (device-generated there)
POWERLAW_N, POWERLAW_NNZ, POWERLAW_ITERS = 4194304, 848535370, 11


def leg_powerlaw(n_sample=262144):
    t0 = time.perf_counter()
    A = O.powerlaw_spd(n_sample, 42)
    t_gen = time.perf_counter() - t0
    b = O.uniform_pm1(42, A.nrows)
    x, s, dt = _pcg(A, b, 1e-30, POWERLAW_ITERS)
    it = s.computed_iterations
    scale = b_pcg(POWERLAW_N, POWERLAW_NNZ) / b_pcg(A.nrows, A.nnz)
    return {"input": f"power-law SPD, n = {n_sample} sample of the "
                     f"n = {POWERLAW_N} workload",
            "solver": f"{it} Jacobi-PCG FP64 iterations",
            "full_solve": False, "extrapolated": True,
            "sample_s": round(dt, 3), "sample_rows": A.nrows,
            "sample_nnz": A.nnz, "generate_s": round(t_gen, 2),
            "s": round(dt * scale, 3),
            "scaling": f"x {scale:.2f} (algorithmic bytes of the full matrix "
                       f"/ the sample's; the sample's x fits in the host "
                       f"caches, so this is optimistic for the CPU)",
            "GBs": round(it * b_pcg(A.nrows, A.nnz) / dt / 1e9, 3)}


def convdiff_peclet(seed=42):
    return 10.0 + 90.0 * ((_mix(seed ^ 0x5045, 0) >> 11) / 2.0 ** 53)


def leg_convdiff(grid=128):
    pe = convdiff_peclet(42)
    h = 1.0 / (grid + 1)
    v = np.array([1.0, 0.5, 0.25])
    c = pe * h * v / np.linalg.norm(v)
    A = O.convdiff_3d(grid, grid, grid, 1.0, *c)
    b = O.uniform_pm1(42, A.nrows)
    t0 = time.perf_counter()
    F = O.ilu0(A)
    t_ilu = time.perf_counter() - t0
    t0 = time.perf_counter()
    x, s = O.gmres(A, b, tol=1e-10, restart_dimension=50, max_iterations=5000,
                   right=("ilu", F), ref_calls=True)
    dt = time.perf_counter() - t0
    return {"input": f"3D upwind convection-diffusion {grid}^3 (Peclet "
                     f"{pe:.3f})", "solver": "FP64 GMRES(50) + exact ILU(0) "
            "to 1e-10", "full_solve": True, "s": round(dt, 3),
            "ilu0_factor_s": round(t_ilu, 3),
            "iterations": s.computed_iterations, "converged": s.converged,
            "true_relres": O.true_relres(A, b, x)}


def leg_batched(count=4096, sample=64, seed=42, smin=32, smax=256):
    sides = np.array([O.batched_side(seed, k, smin, smax) for k in range(count)])
    rows = sides * sides
    offs = np.concatenate([[0], np.cumsum(rows)])
    nnz_all = int(np.sum(5 * rows - 4 * sides))
    bfull = O.uniform_pm1(seed, int(offs[-1]))
    pick = sorted({int(_mix(0xBA7C, j) % count) for j in range(4 * sample)})
    pick = pick[:sample]
    t, its, nnz_s, worst = 0.0, 0, 0, 0.0
    for k in pick:
        A = O.batched_diffusion(seed, k, smin, smax)
        b = bfull[offs[k]:offs[k + 1]]
        x, s, dt = _pcg(A, b, 1e-8, 20000)
        t += dt
        its += s.computed_iterations
        nnz_s += A.nnz
        worst = max(worst, O.true_relres(A, b, x))
    scale = nnz_all / nnz_s
    return {"input": f"{len(pick)} of {count} batched diffusion systems "
                     f"(sides U{{{smin}..{smax}}})",
            "solver": "Jacobi-PCG FP64 to 1e-8 per system",
            "full_solve": False, "extrapolated": True,
            "sample_s": round(t, 3), "sample_systems": len(pick),
            "sample_nnz": nnz_s, "sample_iterations": its,
            "sample_max_true_relres": worst, "s": round(t * scale, 3),
            "scaling": f"x {scale:.2f} (batch nnz / sample nnz)"}


def leg_aniso(n_side=1024):
    A = O.aniso_q1_2d(n_side, 4242, 1.5)
    b = O.uniform_pm1(42, A.nrows)
    x, s, dt = _pcg(A, b, 1e-8, 50000)
    return {"input": f"2D anisotropic Q1 diffusion {n_side}^2",
            "solver": "Jacobi-PCG FP64 to 1e-8", "full_solve": True,
            "s": round(dt, 3), "iterations": s.computed_iterations,
            "converged": s.converged, "true_relres": O.true_relres(A, b, x)}


LEGS = {"p2d": leg_p2d, "poisson192": leg_poisson192,
        "powerlaw": leg_powerlaw, "convdiff": leg_convdiff,
        "batched": leg_batched, "aniso": leg_aniso}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("legs", help="comma-separated: " + ",".join(sorted(LEGS)))
    ap.add_argument("--core", type=int, default=None)
    args = ap.parse_args()
    legs = args.legs.split(",")
    for leg in legs:
        if leg not in LEGS:
            ap.error(f"unknown leg {leg!r}")
    if args.core is not None:
        os.sched_setaffinity(0, {args.core})
    O.build()
    for leg in legs:  # one after another on this core, one line each
        t0 = time.perf_counter()
        r = LEGS[leg]()
        r.update(leg=leg, core=args.core, cores=1, cpu_model=cpu_model(),
                 wall_s=round(time.perf_counter() - t0, 2))
        print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()
