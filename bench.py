#!/usr/bin/env python3
"""Benchmark: FP64 CSR SpMV + Jacobi-PCG on one MI355X.

Workload (the configuration the metric is quoted on): 3D 7-point Poisson,
192^3 grid (7,077,888 rows, 49,324,032 nnz, int32 indices, Dirichlet
diagonal 6), RHS splitmix64(seed 42) uniform[-1,1], x0 = 0, relative
residual tolerance 1e-8, Jacobi left preconditioner, plain PCG
(restart_iterations = max_iterations).  One step = one full PCG solve to
1e-8 with the matrix and vectors already in HBM.

value: algorithmic HBM GB/s sustained over the whole solve
       (iterations x B_pcg / solve time, B_pcg = B_spmv + 11*8*n).
roofline: the SpMV kernel (the dominant kernel), B_spmv per launch / its
       average duration measured with HIP events on the library stream.
cpu_baseline: the oracle's Jacobi-PCG (the CPU restatement of the
       reference) on the same matrix for a bounded number of iterations,
       one core.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def b_spmv(n, nnz, s=8):
    return nnz * (s + 4) + 4 * (n + 1) + 2 * s * n


def cpu_baseline(grid, iters):
    """Oracle Jacobi-PCG, `iters` iterations on the same matrix (1 core)."""
    from oracle import oracle as O
    A = O.poisson_3d(grid)
    b = O.uniform_pm1(42, A.nrows)
    t0 = time.perf_counter()
    _, s = O.cg(A, b, tol=1e-30, restart_iterations=1 << 40,
                max_iterations=iters, left="jacobi")
    dt = time.perf_counter() - t0
    n, nnz = A.nrows, A.nnz
    bytes_ = s.computed_iterations * (b_spmv(n, nnz) + 11 * 8 * n)
    return {"value": round(bytes_ / dt / 1e9, 3), "unit": "GB/s",
            "cores": 1, "kind": "port",
            "sample": f"{s.computed_iterations} Jacobi-PCG iterations of the "
                      f"oracle on the {grid}^3 Poisson matrix "
                      f"({dt:.2f} s; {dt / s.computed_iterations * 1e3:.1f} "
                      f"ms/iteration)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=192)
    ap.add_argument("--spmv-reps", type=int, default=50)
    ap.add_argument("--cpu-iters", type=int, default=12)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(every run uses exactly one MI355X); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import sparkit_amd as spk
    spk.init(0)
    cus, name = spk.device_info()
    log(f"device: {name}, {cus} CUs")
    g = args.grid
    A = spk.poisson_3d(g)
    n, nnz = A.nrows, A.nnz
    log(f"matrix: poisson_3d({g}) n={n} nnz={nnz}")
    b = spk.uniform_pm1(42, n)
    x = spk.DeviceVector.zeros(n)
    J = spk.jacobi(A)
    cfg = spk.CGConfig(tolerance=1e-8, restart_iterations=20000,
                       max_iterations=20000)
    timer = spk.Timer()

    def step():
        call_zero = spk._lib.call
        call_zero("spk_memset", x.ptr, 0, x.nbytes)
        return spk.conjugate_gradient(b, x, cfg, A, J, spk.identity)

    for _ in range(args.warmup):
        s = step()
    spk.synchronize()
    times, iters = [], []
    t_all0 = time.perf_counter()
    for k in range(args.steps):
        timer.start()
        s = step()
        ms = timer.stop()
        times.append(ms)
        iters.append(s.computed_iterations)
        log(f"step {k}: {ms:.2f} ms, {s.computed_iterations} iterations, "
            f"converged={s.converged}")
    spk.synchronize()
    wall = time.perf_counter() - t_all0
    relres = spk.true_relres(A, b, x)
    ms_step = float(np.mean(times))
    bpcg = b_spmv(n, nnz) + 11 * 8 * n
    value = np.mean(iters) * bpcg / (ms_step * 1e-3) / 1e9

    # SpMV kernel alone: back-to-back launches on the library stream
    y = spk.DeviceVector(n)
    for _ in range(5):
        A.spmv(x, y)
    timer.start()
    for _ in range(args.spmv_reps):
        A.spmv(x, y)
    spmv_ms = timer.stop() / args.spmv_reps
    bs = b_spmv(n, nnz)
    spmv_gbs = bs / (spmv_ms * 1e-3) / 1e9
    log(f"spmv: {spmv_ms * 1e3:.1f} us/launch, {spmv_gbs:.0f} GB/s "
        f"({spmv_gbs / HBM_PEAK_GBS:.1%} of 8 TB/s)")

    # device-to-device copy of the same size as one SpMV's traffic: the
    # practical streaming ceiling next to the 8 TB/s peak
    cbytes = (bs // 2) // 8 * 8
    ca, cb = spk.DeviceVector(cbytes // 8), spk.DeviceVector(cbytes // 8)
    spk._lib.call("spk_memset", ca.ptr, 0, cbytes)
    for _ in range(3):
        spk._lib.call("spk_memcpy_dtod", cb.ptr, ca.ptr, cbytes)
    timer.start()
    for _ in range(20):
        spk._lib.call("spk_memcpy_dtod", cb.ptr, ca.ptr, cbytes)
    copy_ms = timer.stop() / 20
    copy_gbs = 2 * cbytes / (copy_ms * 1e-3) / 1e9
    log(f"copy: {copy_gbs:.0f} GB/s (read+write)")

    out = {
        "metric": "time-to-1e-8 residual: learned-FGMRES vs Jacobi-PCG; "
                  "SpMV GB/s % HBM peak",
        "value": round(float(value), 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (device-generated 7-point Poisson, splitmix64 RHS)",
        "config": {"workload": f"FP64 CSR SpMV + Jacobi-PCG to 1e-8, 3D "
                               f"7-point Poisson {g}^3",
                   "rows": n, "nnz": nnz, "index": "int32",
                   "pcg_iterations": int(np.median(iters)),
                   "pcg_time_to_1e-8_ms": round(ms_step, 3),
                   "true_relres": relres,
                   "parallelism": "1 GPU"},
        "spmv": {"us_per_launch": round(spmv_ms * 1e3, 2),
                 "GBs": round(spmv_gbs, 1),
                 "frac_hbm_peak": round(spmv_gbs / HBM_PEAK_GBS, 4),
                 "bytes_per_launch": bs,
                 "frac_of_copy": round(spmv_gbs / copy_gbs, 4)},
        "copy_GBs": round(copy_gbs, 1),
        "roofline": {"bound": "hbm", "achieved": round(spmv_gbs, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(spmv_gbs / HBM_PEAK_GBS, 4),
                     "traffic": None},
        "wall_s_timed": round(wall, 3),
    }
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(g, args.cpu_iters)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
