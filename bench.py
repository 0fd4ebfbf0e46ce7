"""Benchmark: time-steps/s of the TD-BEM hot paths on one MI355X.

Workload (the configuration the headline metric is quoted on that fits one GPU): spherical cavity
of radius 1 in full space, gnomonic cube-sphere 6x16x16 = 1536 bilinear quads (1538 nodes, 4614
DOF), Heaviside internal pressure p = 1, E = rho = 1, nu = 0.25, beta = c_p dt / h = 0.6 with h the
RMS element size (dt = 0.04949, MaxStep 65), 600 steps, alpha = (1/4, 1/4).

One bench "step" is one complete solve of that problem on the GPU: lag-store assembly of all 66
lags, the step-0 factorisations, and the 600-step march (main.cpp:89-128).  The mesh is resident
on the GPU before the timed region.  value = time-steps solved / second over the timed solves.
Printed as one JSON line; --gpus N > 1 exits (one GPU per box in this build).
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "td-bem_amd"))

import numpy as np  # noqa: E402


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--numstep", type=int, default=600, help="time steps per solve")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=16)
    return ap.parse_args()


MI355X_HBM_GBS = 8000.0
MI355X_FP64_TFLOPS = 78.6


def cpu_baseline(p, threads):
    """The oracle (plain-C restatement of the reference, OpenMP over collocation rows) timed on a
    bounded sample of the same workload and scaled to a full solve:
      - assembly: 16 collocation rows at every step 1..MaxStep, scaled by nnd/16;
      - march: the reference's per-step work, i.e. its literal Solve_Averaging, which forms the
        system matrix, refactors it and runs about 6*min(s, MaxStep) dense N x N matvecs per step,
        timed as one LU of an N x N matrix and 8 dense matvecs, scaled over the 600 steps."""
    sys.path.insert(0, str(ROOT / "tests"))
    import ctypes as C
    import oracle_lib as O
    lib = O.oracle()
    K = p.max_step()
    rows = 16
    t0 = time.perf_counter()
    for n in range(1, K + 1):
        O.assemble_step(4, 1, 6, p.E, p.nu, p.rho, p.coords, p.ien, p.if_ext, 0, rows - 1,
                        n * p.dt, p.dt, n, nthreads=threads)
    t_asm = (time.perf_counter() - t0) * p.nnd / rows
    N = p.ndof
    rng = np.random.default_rng(0)
    A = rng.standard_normal((N, N)) + N * np.eye(N)
    piv = np.zeros(N, np.int32)
    t0 = time.perf_counter()
    lib.or_lu_factor(C.c_int(N), A.ctypes.data_as(C.POINTER(C.c_double)),
                     piv.ctypes.data_as(C.POINTER(C.c_int)))
    t_lu = time.perf_counter() - t0
    x = rng.standard_normal(N)
    t0 = time.perf_counter()
    for _ in range(8):
        y = A @ x  # numpy matvec, multithreaded BLAS: a lower bound on the oracle's loop
    t_mv = (time.perf_counter() - t0) / 8
    mv = sum(6 * min(s, K) for s in range(1, p.numstep))
    t_march = (p.numstep - 1) * t_lu + mv * t_mv
    total = t_asm + t_march
    return {"value": p.numstep / total, "unit": "time-steps/s", "cores": threads,
            "kind": "port",
            "sample": (f"oracle assembly of {rows}/{p.nnd} rows x {K} steps ({t_asm * rows / p.nnd:.1f} s"
                       f" measured, x{p.nnd / rows:.0f}); one {N}x{N} LU ({t_lu:.1f} s) per step and"
                       f" {mv} matvecs ({t_mv * 1e3:.1f} ms each) as the literal Solve_Averaging does;"
                       f" extrapolated full solve {total:.0f} s")}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    from tdbem import api, configs

    p = configs.sphere(n=16, numstep=args.numstep)
    ctx = api.Context.from_problem(p)
    ctx.set_profiling(True)

    def one_solve():
        ctx.build_lag_store()
        ctx.march_init(p.a1, p.a2, p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val)
        ctx.march_run()
        ctx.synchronize()

    for _ in range(args.warmup):
        one_solve()
    ctx.reset_timings()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        one_solve()
    wall = time.perf_counter() - t0
    tm = ctx.timings()
    U, T = ctx.history()
    solves = args.steps
    steps_total = solves * (p.numstep - 1)
    value = steps_total / wall
    # dominant kernel: the history convolution (streams the shell lag store once per step)
    conv_avg_ms = tm["conv_ms"] / max(tm["conv_launches"], 1)
    entry_bytes = 15 * 8 + 4
    conv_bytes = tm["store_nonzero"] * entry_bytes
    achieved = conv_bytes / (conv_avg_ms * 1e-3) / 1e9
    out = {
        "metric": "time-steps/s (plain FP64 GPU path, end-to-end assembly + march)",
        "value": round(value, 2),
        "unit": "time-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / solves * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "sphere cavity 6x16x16 (1536 quads, 4614 dof), 600 steps, "
                               "Heaviside p=1, nu=0.25, beta=0.6, plain FP64",
                   "numstep": p.numstep, "max_step": ctx.max_step, "dof": p.ndof},
        "phases_ms_per_solve": {k: round(tm[k] / solves, 3) for k in
                                ("store_ms", "init_ms", "march_ms")},
        "lag_store": {"entries": int(tm["store_nonzero"]), "bytes": ctx.lag_store_bytes(),
                      "dense_bytes": 2 * (ctx.max_step + 1) * p.ndof ** 2 * 8},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": MI355X_HBM_GBS,
                     "unit": "GB/s", "frac": round(achieved / MI355X_HBM_GBS, 4),
                     "traffic": None, "kernel": "k_conv", "avg_ms": round(conv_avg_ms, 4),
                     "bytes_per_launch": int(conv_bytes)},
        "checks": {"u_r_mean_final": float(np.mean(np.einsum(
            "ij,ij->i", U[-1].reshape(-1, 3), p.coords)))},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(p, args.cpu_threads)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
