"""Benchmark: time-steps/s of the TD-BEM hot paths on one MI355X, mixed precision vs plain FP64.

Workload: spherical cavity of radius 1 in full space, gnomonic cube-sphere 6x16x16 = 1536 bilinear
quads (1538 nodes, 4614 DOF), E = rho = 1, nu = 0.25, beta = c_p dt / h = 0.6 with h the RMS
element size (dt = 0.04949, MaxStep 65), 600 steps, alpha = (1/4, 1/4).
  - plain FP64: one Heaviside internal pressure p = 1 (the reference's load);
  - mixed: 32 right-hand sides, pressure amplitudes U[0.5, 2] and rise times U[0, 10 dt] (half of
    them Heaviside), seed 3; split-BF16 MFMA convolution, FP64 assembly and step-0 solve.

One bench "step" is one complete solve: lag-store assembly of all 66 lags, the step-0
factorisations, and the 599 marching solves (src/main.cpp:89-128).  The mesh is resident on the
GPU before the timed region.  value = time-steps solved per second (summed over right-hand sides)
of the mixed form; plain_fp64 gives the same for the plain path (one load, as the reference runs)
and fp64_8rhs for the FP64 path at 8 right-hand sides, its best rate.  The mixed error is the max
over all 32 right-hand sides of ||u_mixed - u_fp64||_F / ||u_fp64||_F (and the same for t)
against FP64 solves of the same loads; above 1e-6 the line is marked invalid and the exit code
is 3.

roofline: the mixed path's dominant kernel k_conv_mixed, HBM-bound: algorithmic bytes per launch =
the in-shell node-pair lags x 18 FP32 values (G's 9 and H's 9 of each 3 x 3 node block) + the
merged-list entries x 32 B + the history copy (split bf16, 6 B a value) read once, over its mean duration from HIP
events on the context stream; stored_bytes_per_launch counts whole stored 16 x 32 sub-tiles
instead.  traffic: HBM bytes per launch measured in this run by two rocprofv3 --pmc passes
(FETCH_SIZE, WRITE_SIZE) over a short run of the same workload (--traffic-probe), each scaled
by what the counter reads for a 1 GiB k_stream_copy; null when rocprofv3 is absent (--no-pmc
skips it).  plain_fp64.roofline gives k_conv the same way (stored entries x 124 B).  assembly:
the lag kernel k_pair_lags against the FP64 vector peak, ASM_FLOPS_PER_POINT flops per active
(quadrature point, lag) evaluation plus ASM_FLOPS_PER_LEVEL per time level evaluated there (the
operation counts of the formulas as the oracle states them; the kernel executes fewer: per-level
constants are hoisted and each level forms only the hats the lag combination needs).
batched and serving: the continuous-batching engine over half-space scenarios and Poisson
request streams at their stated sizes.  Printed as one JSON line; --gpus N > 1 exits (the path
runs on one GPU).
"""
import argparse
import json
import os
import shutil
import signal
import subprocess
import sys
import tempfile
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "td-bem_amd"))

import numpy as np  # noqa: E402

MI355X_HBM_GBS = 8000.0
TP_SUBTILE_BYTES = 16 * 32 * 4           # FP32 16x32 sub-tile
TP_ENTRY_BYTES = 32                      # merged-list step: 4 first sub-tiles, 4 masks, J/k, k-tile mask
FP64_ENTRY_BYTES = 15 * 8 + 4            # 15 values + index per (i, j, k)
MI355X_FP64_TFLOPS = 78.6
# FP64 flops of the lag kernel as written (integrate_points, eval_s, hat_integrals,
# blocks_from_s; +, -, x, / and sqrt 1 each, FMA 2), 4-node elements: per active (point, lag)
# evaluation the point's distance, reciprocals and direction (17), the 15 block values from the
# six time integrals (150) and their accumulation into 4 nodes x 15 values (124); per time level
# evaluated there the two waves' hat integrals (2 x 47) and their six-term accumulation (24).
ASM_FLOPS_PER_POINT = 291
ASM_FLOPS_PER_LEVEL = 118
# rocprofv3 kernel names of the convolution kernels and the calibration copy
PMC_KERNELS = {"k_conv_mixed": "k_conv_mixed(", "k_conv": "k_conv<1>", "copy": "k_stream_copy"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--numstep", type=int, default=600, help="time steps per solve")
    ap.add_argument("--nrhs", type=int, default=32)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-scenario leg")
    # the first scenarios / requests of the seeded stated-size sequences; the whole 64-scenario
    # and 200-request runs take some minutes each (--batch-scenarios 64 --serve-requests 200)
    ap.add_argument("--batch-scenarios", type=int, default=12)
    ap.add_argument("--serve-requests", type=int, default=24)
    ap.add_argument("--serve-rate", type=float, default=2.0, help="Poisson arrivals per second")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 1 GiB device copy (3 launches) to scale HBM counters by")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--fp64-serial-every", type=int, default=2,
                    help="plain FP64 serial baseline of the batched leg on every n-th scenario "
                         "(the serving leg on every 2n-th request); 0 skips it")
    ap.add_argument("--no-pmc", action="store_true",
                    help="skip the two rocprofv3 --pmc passes that measure roofline.traffic")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="(internal) the short run the --pmc passes profile")
    return ap.parse_args()


def cpu_baseline(p, threads):
    """The oracle (plain-C restatement of the reference, OpenMP over collocation rows) timed on a
    bounded sample of the same workload and scaled to a full solve:
      - assembly: 16 collocation rows at every step 1..MaxStep, scaled by nnd/16;
      - march: the reference's per-step work, i.e. its literal Solve_Averaging, which refactors
        the N x N system and runs 6*min(s, MaxStep) dense N x N matvecs per step, timed as one
        oracle LU of an N x N matrix and 8 dense matvecs, scaled over the 599 steps."""
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
    t_rows = time.perf_counter() - t0
    t_asm = t_rows * p.nnd / rows
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
        A @ x
    t_mv = (time.perf_counter() - t0) / 8
    mv = sum(6 * min(s, K) for s in range(1, p.numstep))
    t_march = (p.numstep - 1) * t_lu + mv * t_mv
    total = t_asm + t_march
    return {"value": round((p.numstep - 1) / total, 4), "unit": "time-steps/s", "cores": threads,
            "kind": "port",
            "sample": (f"oracle assembly of {rows}/{p.nnd} rows x {K} steps ({t_rows:.1f} s, "
                       f"scaled x{p.nnd / rows:.0f}); one {N}x{N} oracle LU ({t_lu:.1f} s) per step "
                       f"and {mv} dense matvecs ({t_mv * 1e3:.2f} ms each) as the literal "
                       f"Solve_Averaging does; extrapolated full solve {total:.0f} s")}


def timed_solves(ctx, p, warmup, steps, amp=None, rise=None):
    def one():
        ctx.build_lag_store()
        ctx.march_init(p.a1, p.a2, p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val, amp=amp,
                       t_rise=rise)
        ctx.march_run()
        ctx.synchronize()

    for _ in range(warmup):
        one()
    ctx.reset_timings()
    t0 = time.perf_counter()
    for _ in range(steps):
        one()
    return time.perf_counter() - t0, ctx.timings()


_T0 = time.perf_counter()


def log(msg):
    print(f"[bench {time.perf_counter() - _T0:7.1f}s] {msg}", file=sys.stderr, flush=True)


def copy_calibration(api):
    """k_stream_copy over 1 GiB, three launches: a plain streaming kernel of known size (1 GiB
    read + 1 GiB written per launch) whose FETCH_SIZE / WRITE_SIZE counters scale the counters of
    the kernels above to bytes."""
    ms = api.stream_copy(1 << 30, 3)
    rd = api.stream_read(8 << 30, 3)
    return {"kernel": "k_stream_copy", "bytes_read": 1 << 30, "bytes_written": 1 << 30,
            "launches": 3, "avg_ms": round(ms, 4), "GBs": round(2 * (1 << 30) / ms / 1e6, 1),
            "read_only": {"kernel": "k_stream_read", "bytes": 8 << 30, "launches": 3,
                          "avg_ms": round(rd, 4), "GBs": round((8 << 30) / rd / 1e6, 1)}}


def _engine_summary(res, summary, probs):
    ok = [r for r in res if not r.error]
    out = dict(summary)
    out.update({
        "elements_range": [int(min(p.nel for p in probs)), int(max(p.nel for p in probs))],
        "elements_total": int(sum(p.nel for p in probs)),
        "steps_range": [int(min(p.numstep for p in probs)), int(max(p.numstep for p in probs))],
        "predicted_resident_bytes_total": int(sum(r.predicted_resident for r in ok)),
        "predicted_resident_bytes_max": int(max((r.predicted_resident for r in ok), default=0)),
        "setup_s_total": round(sum(r.setup_s for r in ok), 2),
        "pair_lags_total": int(sum(r.pair_lags for r in ok)),
    })
    return out


def batch_leg(api, configs, serve, count, nrhs=32):
    """Uneven half-space cavity scenarios at the stated sizes (free plane 8x8 + spherical cavity of
    radius U[0.5, 1.5] at depth U[2, 4]; elements U[400, 4000], steps U[100, 800], nu U[0.2,
    0.35], seed 4), nrhs pressure histories each (32, as the sphere workload), all arriving at
    once: the continuous-batching
    engine admits them by their predicted device bytes, sets them up 4 at a time while the
    admitted ones march together in one batch (hipGraph replays of 8 steps of every member), and
    lets the next ones in as finished ones leave.  Time-steps/s over the whole job, assembly
    included."""
    scen = configs.halfspace_scenarios(count=count, seed=4)
    reqs = serve.poisson_requests(scen, rate=None, nrhs=nrhs, seed=40)
    res, summary = serve.Engine(nrhs=nrhs, setup_workers=4, steps_per_graph=8,
                                tune=True, progress=log).run(reqs)
    out = _engine_summary(res, summary, scen)
    out.update({"scenarios": count, "nrhs": nrhs})
    return out


def serve_leg(api, configs, serve, nreq, rate, nrhs=32):
    """Serving at the stated sizes: Poisson arrivals (rate per second, seed 5) of solves of the
    bar and sphere geometries, element counts log-uniform in [300, 6000], steps U[100, 1000],
    32 pressure histories each, through the continuous-batching engine (admission by predicted
    device bytes against the device's free memory, set-ups beside the batched march)."""
    probs = configs.serving_mix(nreq, seed=5)
    reqs = serve.poisson_requests(probs, rate=rate, nrhs=nrhs, seed=5)
    res, summary = serve.Engine(nrhs=nrhs, setup_workers=4, steps_per_graph=8,
                                tune=True, progress=log).run(reqs)
    out = _engine_summary(res, summary, probs)
    out.update({"arrival_rate_per_s": rate, "nrhs": nrhs})
    return out


def fp64_serial(api, probs, every, steps=8):
    """The plain FP64 path (one load per solve, as the reference runs) on the leg's own requests,
    one after another on an idle device: requests 0, every, 2 every, ... each set up (FP64 lag
    store, step-0 factorisations) and marched for 2 + `steps` steps, the per-step time of the
    last `steps` (every step past the first streams the whole store, so it is constant) carried to
    the request's step count.  time_steps_per_s = sum of (numstep - 1) / sum of the estimated
    solve times."""
    rows = []
    for i in range(0, len(probs), every):
        p = probs[i]
        with api.Context.from_problem(p) as ctx:
            t0 = time.perf_counter()
            ctx.build_lag_store()
            ctx.march_init(p.a1, p.a2, p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val)
            ctx.synchronize()
            t1 = time.perf_counter()
            ctx.march_run(2)
            ctx.synchronize()
            t2 = time.perf_counter()
            n = min(steps, p.numstep - 3)
            ctx.march_run(n)
            ctx.synchronize()
            t3 = time.perf_counter()
        step = (t3 - t2) / max(n, 1)
        rows.append({"request": i, "elements": int(p.nel), "steps": int(p.numstep),
                     "setup_s": round(t1 - t0, 3), "step_ms": round(step * 1e3, 3),
                     "solve_s": round((t2 - t0) + (p.numstep - 3) * step, 3)})
    total = sum(r["solve_s"] for r in rows)
    tsteps = sum(r["steps"] - 1 for r in rows)
    return {"requests_measured": len(rows), "every": every, "steps_timed": steps,
            "time_steps_per_s": round(tsteps / total, 2) if total > 0 else 0.0,
            "serial_solve_s": round(total, 2), "per_request": rows}


def traffic_probe(nrhs):
    """What the counter passes profile: the 1 GiB calibration copy, then one mixed solve at nrhs
    right-hand sides and one plain FP64 solve of the bench workload, 80 steps (every launch past
    step MaxStep streams the whole store, as at 600 steps)."""
    from tdbem import api, configs
    api.stream_copy(1 << 30, 2)
    p = configs.sphere(n=16, numstep=80)
    amp, rise = configs.rhs_set(nrhs, p.dt, seed=3)
    with api.Context.from_problem(p) as ctx:
        ctx.set_precision(api.PRECISION_MIXED)
        timed_solves(ctx, p, 0, 1, amp, rise)
    with api.Context.from_problem(p) as ctx:
        timed_solves(ctx, p, 0, 1)


def measure_traffic(nrhs):
    """HBM bytes per launch of k_conv_mixed and k_conv: two rocprofv3 --pmc passes (FETCH_SIZE,
    then WRITE_SIZE, KiB per dispatch) over traffic_probe, each counter scaled by its reading for
    the 1 GiB k_stream_copy (1 GiB read, 1 GiB written).  The mean over the last 10 launches of
    each kernel.  Returns ({kernel: bytes}, note)."""
    exe = shutil.which("rocprofv3")
    if not exe:
        return {}, "rocprofv3 not found: traffic not measured"
    kib = {}
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            out_dir = os.path.join(d, counter)
            cmd = [exe, "--pmc", counter, "--kernel-include-regex", "k_conv|k_stream_copy",
                   "-f", "csv", "-d", out_dir, "-o", "pmc", "--", sys.executable, "-u",
                   str(ROOT / "bench.py"), "--traffic-probe", "--nrhs", str(nrhs)]
            env = dict(os.environ, TMPDIR="/tmp")
            proc = subprocess.Popen(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL,
                                    stderr=subprocess.DEVNULL, start_new_session=True)
            try:
                rc = proc.wait(timeout=240)
            except subprocess.TimeoutExpired:
                os.killpg(proc.pid, signal.SIGKILL)
                proc.wait()
                return {}, f"rocprofv3 --pmc {counter} pass timed out"
            csvs = list(Path(out_dir).rglob("*counter_collection.csv"))
            if not csvs:
                return {}, f"rocprofv3 --pmc {counter} pass wrote no counters (exit {rc})"
            import csv
            vals = {k: [] for k in PMC_KERNELS}
            with open(csvs[0]) as f:
                for row in csv.DictReader(f):
                    if row["Counter_Name"] != counter:
                        continue
                    for k, pat in PMC_KERNELS.items():
                        if pat in row["Kernel_Name"]:
                            vals[k].append(float(row["Counter_Value"]))
            if not vals["copy"]:
                return {}, f"rocprofv3 --pmc {counter}: no calibration dispatch"
            kib[counter] = {k: (sum(v[-10:]) / len(v[-10:]) if v else None)
                            for k, v in vals.items()}
    f_read = (1 << 30) / (kib["FETCH_SIZE"]["copy"] * 1024)
    f_write = (1 << 30) / (kib["WRITE_SIZE"]["copy"] * 1024)
    res = {}
    for k in ("k_conv_mixed", "k_conv"):
        rd, wr = kib["FETCH_SIZE"][k], kib["WRITE_SIZE"][k]
        if rd is not None and wr is not None:
            res[k] = {"bytes": rd * 1024 * f_read + wr * 1024 * f_write,
                      "read": rd * 1024 * f_read, "written": wr * 1024 * f_write}
    note = (f"rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes over --traffic-probe in this run; "
            f"scale factors {f_read:.4f} / {f_write:.4f} from a 1 GiB k_stream_copy")
    return res, note


def roofline(kernel, avg_ms, nbytes, traffic=None):
    ach = nbytes / (avg_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(ach, 1), "peak": MI355X_HBM_GBS, "unit": "GB/s",
            "frac": round(ach / MI355X_HBM_GBS, 4), "traffic": traffic, "kernel": kernel,
            "avg_ms": round(avg_ms, 4), "bytes_per_launch": int(nbytes)}


def main():
    args = parse()
    if args.traffic_probe:
        traffic_probe(args.nrhs)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "the path runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    from tdbem import api, configs

    p = configs.sphere(n=16, numstep=args.numstep)
    R = args.nrhs
    amp, rise = configs.rhs_set(R, p.dt, seed=3)
    K = p.max_step()
    out = {"metric": "time-steps/s mixed vs plain FP64 GPU at <=1e-6 err",
           "unit": "time-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "bf16x3-mfma/f64", "data": "synthetic",
           "config": {"workload": f"sphere cavity 6x16x16 (1536 quads, 4614 dof), {p.numstep} "
                                  f"steps, nu=0.25, beta=0.6; mixed: {R} pressure histories per "
                                  f"solve", "numstep": p.numstep, "max_step": K, "dof": p.ndof,
                      "nrhs": R}}

    if not args.no_mixed:
        log(f"mixed leg: {R} right-hand sides, {p.numstep} steps")
        with api.Context.from_problem(p) as ctx:
            # timed solves: events around the convolution only (the roofline's launch time);
            # then one untimed solve with every kernel timed for the reduction and step solve
            ctx.set_profiling(True, conv_only=True)
            ctx.set_precision(api.PRECISION_MIXED)
            wall, tm = timed_solves(ctx, p, args.warmup, args.steps, amp, rise)
            mixed_hist = [ctx.history(r) for r in range(R)]
            ctx.set_profiling(True)
            _, tm_all = timed_solves(ctx, p, 0, 1, amp, rise)
        nconv = max(tm["conv_launches"], 1)
        conv_ms = tm["conv_ms"] / nconv
        # per launch (time-blocked: the near lags k < 3 for step s and the far lags of
        # parity s mod 2 for steps s and s + 1, so a launch streams the near lags and half the
        # far ones; the mean of the two parities), the algorithmic bytes: the in-shell node-pair
        # lags it reads (18 FP32 values each, G's 9 and H's 9, as the 48 x 96 blocks hold them),
        # its merged lists, and the history copy read once ((K + 1) slots x 96 columns per
        # tile x R x 6 B, three bf16 terms a value)
        blocked = True
        nT = -(-p.nnd // 16)
        hist_bytes = (K + 1) * nT * 96 * R * 6
        if blocked:
            lags_launch = tm["pair_lags_near"] + 0.5 * (tm["pair_lags_even"] + tm["pair_lags_odd"])
            subs_launch = tm["subtiles_near"] + 0.5 * (tm["subtiles_even"] + tm["subtiles_odd"])
            ents_launch = 0.5 * (tm["entries_parity0"] + tm["entries_parity1"])
        else:
            lags_launch, subs_launch = tm["store_nonzero"], tm["tile_subtiles"]
            ents_launch = tm["tile_entries"]
        list_bytes = ents_launch * TP_ENTRY_BYTES
        inshell_bytes = lags_launch * 18 * 4 + list_bytes + hist_bytes
        # what the kernel streams: whole stored 16 x 32 sub-tiles, zeros inside them included
        stored_bytes = subs_launch * TP_SUBTILE_BYTES + list_bytes + hist_bytes
        # one step's convolution as an unblocked kernel reads it: every in-shell node-pair lag
        step_bytes = (tm["store_nonzero"] * 18 * 4 + tm["tile_entries"] * TP_ENTRY_BYTES +
                      hist_bytes)
        out.update({
            "value": round(args.steps * (p.numstep - 1) * R / wall, 2),
            "ms_per_step": round(wall / args.steps * 1e3, 3),
            "mixed": {"phases_ms_per_solve": {k: round(tm[k] / args.steps, 3)
                                              for k in ("store_ms", "init_ms", "march_ms")},
                      "pair_lags": int(tm["store_nonzero"]),
                      "tile_subtiles": int(tm["tile_subtiles"]),
                      "tile_blocks": int(tm["tile_blocks"]),
                      "tile_entries": int(tm["tile_entries"]),
                      "tile_store_bytes": int(tm["tile_subtiles"] * TP_SUBTILE_BYTES),
                      "subtile_padding": round(tm["tile_subtiles"] * TP_SUBTILE_BYTES /
                                               max(tm["store_nonzero"] * 18 * 4, 1), 3),
                      "conv_avg_ms": round(conv_ms, 4),
                      "reduce_avg_ms": round(tm_all["reduce_ms"] / max(tm_all["conv_launches"], 1), 4),
                      "solve_avg_ms": round(tm_all["solve_ms"] / max(tm_all["conv_launches"], 1), 4),
                      "lag_kernel_ms": round(tm_all["lag_kernel_ms"], 3)},
        })
        # the lag kernel of the same store build (one solve, every kernel timed)
        asm_flops = (tm_all["active_points"] * ASM_FLOPS_PER_POINT +
                     tm_all["active_levels"] * ASM_FLOPS_PER_LEVEL)
        asm_tf = asm_flops / (tm_all["lag_kernel_ms"] * 1e-3) / 1e12
        out["assembly"] = {"bound": "fp64-valu", "kernel": "k_pair_lags (lag-store assembly)",
                           "achieved": round(asm_tf, 2), "peak": MI355X_FP64_TFLOPS,
                           "unit": "TFLOP/s", "frac": round(asm_tf / MI355X_FP64_TFLOPS, 4),
                           "kernel_ms_per_store": round(tm_all["lag_kernel_ms"], 3),
                           "active_points": int(tm_all["active_points"]),
                           "active_levels": int(tm_all["active_levels"]),
                           "flops_per_store": int(asm_flops)}
        out["roofline"] = roofline("k_conv_mixed (split-BF16 MFMA history convolution)", conv_ms,
                                   inshell_bytes)
        out["roofline"]["stored_bytes_per_launch"] = int(stored_bytes)
        out["roofline"]["frac_of_stored_bytes"] = round(
            stored_bytes / (conv_ms * 1e-3) / 1e9 / MI355X_HBM_GBS, 4)
        out["roofline"]["time_blocked"] = blocked
        # one step's in-shell bytes (the whole store, as an unblocked kernel streams it every
        # step) over the time per step: the rate an unblocked kernel would need to keep pace
        out["roofline"]["step_bytes"] = int(step_bytes)
        out["roofline"]["step_equivalent_GBs"] = round(step_bytes / (conv_ms * 1e-3) / 1e9, 1)
        out["roofline"]["step_equivalent_frac"] = round(
            step_bytes / (conv_ms * 1e-3) / 1e9 / MI355X_HBM_GBS, 4)

        # FP64 at 8 right-hand sides: the FP64 path's best rate (timed), then the error reference
        # for every one of the 32 right-hand sides (four untimed 8-RHS solves in all)
        log("FP64 8-RHS leg (timed, then the error reference for all right-hand sides)")
        n8 = min(args.steps, 3)
        ref = []
        with api.Context.from_problem(p) as ctx:
            ctx.set_profiling(True, conv_only=True)
            wall8, tm8 = timed_solves(ctx, p, min(args.warmup, 1), n8, amp[:8], rise[:8])
            ref += [ctx.history(r) for r in range(8)]
            ctx.set_profiling(False)
            for r0 in range(8, R, 8):
                ctx.march_init(p.a1, p.a2, p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val,
                               amp=amp[r0:r0 + 8], t_rise=rise[r0:r0 + 8])
                ctx.march_run()
                ref += [ctx.history(r) for r in range(8)]
        errs = [max(np.linalg.norm(m[0] - r[0]) / np.linalg.norm(r[0]),
                    np.linalg.norm(m[1] - r[1]) / np.linalg.norm(r[1]))
                for m, r in zip(mixed_hist, ref)]
        out["mixed"]["max_rel_err_vs_fp64"] = float(max(errs))
        out["mixed"]["rhs_checked"] = len(errs)
        out["mixed"]["rel_err_per_rhs"] = [float(f"{e:.3e}") for e in errs]
        out["fp64_8rhs"] = {"value": round(n8 * (p.numstep - 1) * 8 / wall8, 2),
                            "solves": n8,
                            "ms_per_step": round(wall8 / n8 * 1e3, 3),
                            "conv_avg_ms": round(tm8["conv_ms"] / max(tm8["conv_launches"], 1), 4)}
        if len(errs) != R or max(errs) > 1e-6:
            # the headline is stated at <= 1e-6 error: a larger error invalidates it
            out["valid"] = False
            out["invalid_reason"] = f"mixed error {max(errs):.3e} > 1e-6 or RHS unchecked"

    if not args.no_fp64:
        log("plain FP64 leg")
        with api.Context.from_problem(p) as ctx:
            ctx.set_profiling(True, conv_only=True)
            n64 = min(args.steps, 5)
            wall64, tm64 = timed_solves(ctx, p, min(args.warmup, 1), n64)
            U, _ = ctx.history()
            store_bytes = ctx.lag_store_bytes()
            ctx.set_profiling(True)   # one more solve, untimed: the lag kernel and its points
            _, tm64_all = timed_solves(ctx, p, 0, 1)
        conv64 = tm64["conv_ms"] / max(tm64["conv_launches"], 1)
        # per launch: every stored (i, j, k) entry (15 FP64 values and a 4 B index)
        bytes64 = tm64["store_nonzero"] * FP64_ENTRY_BYTES
        out["plain_fp64"] = {
            "value": round(n64 * (p.numstep - 1) / wall64, 2),
            "solves": n64,
            "ms_per_step": round(wall64 / n64 * 1e3, 3),
            "phases_ms_per_solve": {k: round(tm64[k] / n64, 3)
                                    for k in ("store_ms", "init_ms", "march_ms")},
            "lag_kernel_ms": round(tm64_all["lag_kernel_ms"], 3),
            "active_points_per_solve": int(tm64_all["active_points"]),
            "store": {"entries": int(tm64["store_nonzero"]), "bytes": store_bytes,
                      "dense_bytes": 2 * (K + 1) * p.ndof ** 2 * 8},
            "roofline": roofline("k_conv (plain FP64 history convolution)", conv64, bytes64),
            # Lame static radial displacement p a / (4 mu) = 0.625 at late time
            "u_r_mean_final": float(np.mean(np.einsum("ij,ij->i", U[-1].reshape(-1, 3),
                                                      p.coords))),
        }
        if "value" in out:
            out["speedup_mixed_vs_plain_fp64"] = round(out["value"] / out["plain_fp64"]["value"], 2)
            if "fp64_8rhs" in out:
                out["speedup_mixed_vs_fp64_8rhs"] = round(out["value"] / out["fp64_8rhs"]["value"], 2)
        else:
            out["value"] = out["plain_fp64"]["value"]
            out["ms_per_step"] = out["plain_fp64"]["ms_per_step"]
            out["dtype"] = "f64"
            out["roofline"] = out["plain_fp64"]["roofline"]
    from tdbem import serve
    if not args.no_batch:
        log(f"batched leg: {args.batch_scenarios} half-space scenarios at the stated sizes")
        out["batched"] = batch_leg(api, configs, serve, args.batch_scenarios)
        if args.fp64_serial_every > 0:
            log("plain FP64 serial solves of the batched leg's scenarios")
            scen = configs.halfspace_scenarios(count=args.batch_scenarios, seed=4)
            ser = fp64_serial(api, scen, args.fp64_serial_every)
            out["batched"]["plain_fp64_serial"] = ser
            out["batched"]["speedup_vs_plain_fp64_serial"] = round(
                out["batched"]["time_steps_per_s"] / ser["time_steps_per_s"], 2)
    if not args.no_serve:
        log(f"serving leg: {args.serve_requests} requests at {args.serve_rate}/s")
        out["serving"] = serve_leg(api, configs, serve, args.serve_requests, args.serve_rate)
        if args.fp64_serial_every > 0:
            log("plain FP64 serial solves of the serving leg's requests")
            ser = fp64_serial(api, configs.serving_mix(args.serve_requests, seed=5),
                              2 * args.fp64_serial_every)
            out["serving"]["plain_fp64_serial"] = ser
            out["serving"]["speedup_vs_plain_fp64_serial"] = round(
                out["serving"]["time_steps_per_s"] / ser["time_steps_per_s"], 2)
    if args.copy_calibration:
        log("copy calibration kernel")
        out["copy_calibration"] = copy_calibration(api)
    if not args.no_pmc and ("roofline" in out):
        log("HBM traffic: two rocprofv3 --pmc passes over a short run")
        traffic, note = measure_traffic(R)
        for key, k in (("roofline", "k_conv_mixed"), ("plain_fp64", "k_conv")):
            rl = out[key] if key == "roofline" else out.get(key, {}).get("roofline")
            if rl is None or k not in rl.get("kernel", ""):
                continue
            rl["traffic"] = round(traffic[k]["bytes"]) if k in traffic else None
            rl["traffic_source"] = note
            if k in traffic:
                rl["traffic_read"] = round(traffic[k]["read"])
                rl["traffic_written"] = round(traffic[k]["written"])
                rl["traffic_over_algorithmic"] = round(traffic[k]["bytes"] / rl["bytes_per_launch"], 3)
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle sample)")
        out["cpu_baseline"] = cpu_baseline(p, args.cpu_threads)
    print(json.dumps(out), flush=True)
    if out.get("valid") is False:
        sys.exit(3)


if __name__ == "__main__":
    main()
