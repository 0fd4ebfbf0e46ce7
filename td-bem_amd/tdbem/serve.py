"""Serving loop: time-marching solves admitted by the HBM their lag stores need, kept resident on
the GPU and run concurrently (continuous batching), with p50/p99 latency beside throughput.

Each admitted request gets its own context (own HIP stream, own device state).  A worker thread
per request drives assembly, the step-0 factorisations and the march through the C ABI; ctypes
drops the GIL inside every library call, so the requests' kernels overlap on the device while the
admission loop keeps accepting arrivals.  Admission is FIFO: a request starts when an execution
slot is free and its predicted footprint fits the remaining HBM budget.
"""
import threading
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import api

HBM_BYTES = 288e9


@dataclass
class Request:
    rid: int
    problem: object            # configs.Problem
    arrival: float             # seconds after the start of the run
    amp: np.ndarray            # per right-hand side
    rise: np.ndarray
    precision: int = api.PRECISION_MIXED


@dataclass
class Result:
    rid: int
    arrival: float
    admitted: float = 0.0
    finished: float = 0.0
    time_steps: int = 0
    predicted_bytes: int = 0
    error: str = ""
    histories: list = field(default_factory=list)

    @property
    def latency(self):
        return self.finished - self.arrival


def predict_bytes(ctx, nrhs, precision, numstep):
    """Peak device bytes of one request: shell store (FP64, always built), its tile-pair copy for
    the mixed path, the dense step-0 work set during factorisation, and the march state."""
    store = ctx.lag_store_bytes()
    N = ctx.N
    tp = int(1.6 * store) if precision == api.PRECISION_MIXED else 0
    dense = 8 * N * N * 8
    march = (2 * numstep + 140) * nrhs * N * 8
    return store + tp + dense + march


def _run_one(ctx, req, res, keep, done_cb):
    try:
        p = req.problem
        ctx.set_precision(req.precision)
        ctx.build_lag_store()
        ctx.march_init(p.a1, p.a2, p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val, amp=req.amp,
                       t_rise=req.rise)
        ctx.march_run()
        res.time_steps = (p.numstep - 1) * len(req.amp)
        if keep:
            res.histories = [ctx.history(r) for r in range(len(req.amp))]
    except Exception as e:  # a failing request is reported, the others go on
        res.error = str(e)
    finally:
        ctx.close()
        done_cb(req, res)


class Server:
    def __init__(self, hbm_budget=0.9 * HBM_BYTES, max_active=8, keep_histories=False):
        self.hbm_budget = hbm_budget
        self.max_active = max_active
        self.keep = keep_histories

    def run(self, requests, poll=0.002):
        """Runs the requests against their arrival times; returns the results in request order
        and a summary (throughput over the makespan, latency percentiles)."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        waiting = deque()
        results = {r.rid: Result(r.rid, r.arrival) for r in requests}
        lock = threading.Lock()
        state = {"used": 0, "active": 0}
        threads = []
        t0 = time.perf_counter()

        def done(req, res):
            with lock:
                res.finished = time.perf_counter() - t0
                state["used"] -= res.predicted_bytes
                state["active"] -= 1

        while pending or waiting or state["active"] > 0:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            while waiting:
                req = waiting[0]
                with lock:
                    if state["active"] >= self.max_active:
                        break
                ctx = api.Context.from_problem(req.problem)
                need = predict_bytes(ctx, len(req.amp), req.precision, req.problem.numstep)
                with lock:
                    fits = state["used"] + need <= self.hbm_budget or state["active"] == 0
                    if fits:
                        state["used"] += need
                        state["active"] += 1
                if not fits:
                    ctx.close()
                    break
                waiting.popleft()
                res = results[req.rid]
                res.admitted = time.perf_counter() - t0
                res.predicted_bytes = need
                th = threading.Thread(target=_run_one, args=(ctx, req, res, self.keep, done),
                                      daemon=True)
                th.start()
                threads.append(th)
            time.sleep(poll)
        for th in threads:
            th.join()
        makespan = time.perf_counter() - t0
        res = [results[r.rid] for r in requests]
        return res, summarize(res, makespan)


def summarize(results, makespan):
    ok = [r for r in results if not r.error]
    lat = np.array([r.latency for r in ok]) if ok else np.zeros(1)
    steps = sum(r.time_steps for r in ok)
    return {"requests": len(results), "failed": len(results) - len(ok),
            "time_steps": int(steps), "makespan_s": round(makespan, 3),
            "time_steps_per_s": round(steps / makespan, 2) if makespan > 0 else 0.0,
            "latency_p50_s": round(float(np.percentile(lat, 50)), 3),
            "latency_p99_s": round(float(np.percentile(lat, 99)), 3)}


def poisson_requests(problems, rate, nrhs=None, seed=5, precision=api.PRECISION_MIXED):
    """Requests over the given problems with exponential inter-arrival times (rate per second),
    per-request load amplitudes U[0.5, 2] and rise times U[0, 10 dt]; nrhs None takes each
    problem's meta["nrhs"]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for i, p in enumerate(problems):
        t += rng.exponential(1.0 / rate)
        n = nrhs if nrhs is not None else int(p.meta.get("nrhs", 16))
        amp = rng.uniform(0.5, 2.0, n)
        rise = np.where(rng.uniform(size=n) < 0.5, 0.0, rng.uniform(0, 10 * p.dt, n))
        reqs.append(Request(i, p, t, amp, rise, precision))
    return reqs
