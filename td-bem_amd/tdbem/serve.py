"""Continuous batching of time-marching solves on one GPU, with p50/p99 latency beside throughput.

Every admitted solve is stepped in the same kernel launches: the engine keeps one batch
(tdbem_batch_*) whose members are the running solves, and replays a hipGraph of a few steps of
all of them (six launches per step, whatever the number of members).  Around the batch:

  - arrivals join a FIFO queue at their arrival times;
  - each queued request's device footprint is predicted off the main loop (tdbem_predict_bytes:
    the store planner's exact counts of node-pair lags and stored sub-tiles, the march state and
    the set-up's temporaries), and the head of the queue is admitted when a set-up worker is free
    and its predicted peak fits the HBM budget (or nothing else is on the device);
  - an admitted request's set-up (lag and tile store, step-0 factorisation and inverse) runs on a
    worker thread, on its own context's stream, while the batch marches;
  - a finished set-up joins the batch at the next graph boundary; members past their last step
    leave it, their histories are read and their memory is released, which admits the next.

All requests of one engine have the same number of right-hand sides (one batch).  A solve's
results do not depend on which other solves share its launches (each scenario's convolution is
cut into chunks by its own work only), so a request's histories are the same alone and in any mix.
"""
import time
from collections import deque
from concurrent.futures import FIRST_COMPLETED, ThreadPoolExecutor, wait
from dataclasses import dataclass, field

import numpy as np

from . import api

HBM_BYTES = 288e9
# freed set-up memory the device's memory pool may keep (tdbem's pool release threshold)
POOL_KEEP_BYTES = 16 << 30


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
    admitted: float = 0.0      # set-up started
    joined: float = 0.0        # joined the batch
    finished: float = 0.0
    time_steps: int = 0
    predicted_bytes: int = 0   # peak
    predicted_resident: int = 0
    nel: int = 0
    numstep: int = 0
    pair_lags: int = 0         # in-shell (row node, column node, lag) entries of its store
    setup_s: float = 0.0       # set-up wall time (lag and tile store, step-0 solve)
    error: str = ""
    histories: list = field(default_factory=list)

    @property
    def latency(self):
        return self.finished - self.arrival


def _predict(req, nrhs):
    """A context for the request and its predicted (resident, peak) bytes (runs on a planner
    thread: context creation and the store planning are C calls that release the GIL)."""
    ctx = api.Context.from_problem(req.problem)
    try:
        ctx.set_precision(req.precision)
        resident, peak = ctx.predict_bytes(nrhs)
    except BaseException:
        ctx.close()
        raise
    return ctx, resident, peak


def _setup(req, ctx):
    """Set-up of an admitted request on its context; returns its wall time."""
    p = req.problem
    t0 = time.perf_counter()
    ctx.build_lag_store()
    ctx.march_init(p.a1, p.a2, p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val, amp=req.amp,
                   t_rise=req.rise)
    return time.perf_counter() - t0


class Engine:
    """Continuous-batching engine (see the module docstring).

    nrhs: right-hand sides of every request; hbm_budget: bytes the engine may commit (default:
    the device's free memory at the start less POOL_KEEP_BYTES and a 4 GiB margin);
    setup_workers: set-ups in flight at once; steps_per_graph: steps of all members per graph
    replay (the granularity at which members join and leave); tune: time the batch's tunables
    once, when it first holds two members (else the defaults, which the single-context path
    uses too); keep_histories: read every member's u and t histories when it finishes;
    progress: called with a line of text as each request finishes or fails."""

    def __init__(self, nrhs=32, hbm_budget=None, setup_workers=4, steps_per_graph=8,
                 tune=False, keep_histories=False, poll=0.002, progress=None):
        self.nrhs = int(nrhs)
        self.hbm_budget = hbm_budget
        self.setup_workers = max(1, int(setup_workers))
        self.steps_per_graph = int(steps_per_graph)
        self.tune = tune
        self.keep = keep_histories
        self.poll = poll
        self.progress = progress

    def run(self, requests):
        """Serves the requests against their arrival times; returns the results in request
        order and a summary (throughput over the makespan, latency percentiles, batch use)."""
        for r in requests:
            if len(r.amp) != self.nrhs:
                raise ValueError(f"request {r.rid}: {len(r.amp)} right-hand sides, engine "
                                 f"has {self.nrhs}")
        budget = self.hbm_budget
        if budget is None:
            free, _ = api.device_memory()
            budget = free - POOL_KEEP_BYTES - (4 << 30)
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        waiting = deque()
        results = {r.rid: Result(r.rid, r.arrival, nel=int(r.problem.nel),
                                 numstep=int(r.problem.numstep)) for r in requests}
        planned = {}                   # rid -> future of (ctx, resident, peak)
        setups = {}                    # future -> (req, ctx, peak)
        members = {}                   # ctx -> (req, peak)
        used = 0
        stats = {"graph_launches": 0, "member_steps": 0, "max_members": 0, "tuned": None,
                 "budget_bytes": int(budget), "max_committed_bytes": 0}
        batch = api.Batch()
        planner = ThreadPoolExecutor(max_workers=2)
        workers = ThreadPoolExecutor(max_workers=self.setup_workers)
        t0 = time.perf_counter()
        clock = lambda: time.perf_counter() - t0  # noqa: E731

        done = [0]

        def report(req, what):
            done[0] += 1
            if self.progress:
                self.progress(f"request {req.rid} {what} ({done[0]}/{len(requests)} done, "
                              f"{len(members)} in the batch)")

        def fail(req, ctx, msg):
            res = results[req.rid]
            res.error = msg
            res.finished = clock()
            report(req, f"failed: {msg}")
            if ctx is not None:
                try:
                    ctx.close()
                except Exception:   # noqa: BLE001 - reported through res.error already
                    pass

        try:
            while pending or waiting or setups or members:
                now = clock()
                while pending and pending[0].arrival <= now:
                    req = pending.popleft()
                    waiting.append(req)
                    planned[req.rid] = planner.submit(_predict, req, self.nrhs)
                # admission: FIFO, by predicted peak bytes
                while waiting and len(setups) < self.setup_workers:
                    req = waiting[0]
                    fut = planned[req.rid]
                    if not fut.done():
                        break
                    try:
                        ctx, resident, peak = fut.result()
                    except Exception as e:  # noqa: BLE001 - a failing request is reported
                        waiting.popleft()
                        fail(req, None, str(e))
                        continue
                    if used + peak > budget and (members or setups):
                        break
                    waiting.popleft()
                    used += peak
                    stats["max_committed_bytes"] = max(stats["max_committed_bytes"], used)
                    res = results[req.rid]
                    res.admitted = clock()
                    res.predicted_bytes = int(peak)
                    res.predicted_resident = int(resident)
                    setups[workers.submit(_setup, req, ctx)] = (req, ctx, peak)
                # finished set-ups join the batch
                for f in [f for f in setups if f.done()]:
                    req, ctx, peak = setups.pop(f)
                    try:
                        results[req.rid].setup_s = f.result()
                        results[req.rid].pair_lags = int(ctx.timings()["store_nonzero"])
                        batch.add(ctx)
                    except Exception as e:  # noqa: BLE001
                        used -= peak
                        fail(req, ctx, str(e))
                        continue
                    results[req.rid].joined = clock()
                    members[ctx] = (req, peak)
                if members:
                    stats["max_members"] = max(stats["max_members"], len(members))
                    if self.tune and stats["tuned"] is None and len(members) >= 2:
                        stats["tuned"] = batch.tune()
                    before = {c: c.step_index for c in members}
                    stats["graph_launches"] += batch.run(steps_per_graph=self.steps_per_graph,
                                                         max_replays=1)
                    now = clock()
                    for ctx in list(members):
                        stats["member_steps"] += ctx.step_index - before[ctx]
                        req, peak = members[ctx]
                        if ctx.step_index < req.problem.numstep - 1:
                            continue
                        batch.remove(ctx)
                        del members[ctx]
                        res = results[req.rid]
                        res.finished = now
                        res.time_steps = (req.problem.numstep - 1) * self.nrhs
                        if self.keep:
                            res.histories = [ctx.history(r) for r in range(self.nrhs)]
                        ctx.close()
                        used -= peak
                        report(req, f"finished, {req.problem.nel} elements x "
                                    f"{req.problem.numstep} steps")
                elif setups:
                    wait(list(setups), timeout=0.05, return_when=FIRST_COMPLETED)
                elif waiting:
                    wait([planned[waiting[0].rid]], timeout=0.05)
                elif pending:
                    time.sleep(max(0.0, min(self.poll, pending[0].arrival - clock())))
        finally:
            # after an error: wait for the set-ups in flight, then free every context still held
            workers.shutdown(wait=True)
            planner.shutdown(wait=True)
            for req, ctx, _ in setups.values():
                fail(req, ctx, "engine stopped")
            for ctx in list(members):
                batch.remove(ctx)
                fail(members[ctx][0], ctx, "engine stopped")
            batch.close()
            for req in waiting:
                f = planned[req.rid]
                if f.exception() is None:
                    fail(req, f.result()[0], "engine stopped")
        makespan = clock()
        res = [results[r.rid] for r in requests]
        summary = summarize(res, makespan)
        launches = max(stats["graph_launches"], 1)
        summary.update({"engine_graph_launches": stats["graph_launches"],
                        "engine_max_members": stats["max_members"],
                        "engine_mean_members_per_step": round(
                            stats["member_steps"] / (launches * self.steps_per_graph), 2),
                        "engine_budget_bytes": stats["budget_bytes"],
                        "engine_max_committed_bytes": stats["max_committed_bytes"]})
        if stats["tuned"] is not None:
            summary["engine_tuned"] = stats["tuned"]
        return res, summary


def summarize(results, makespan):
    ok = [r for r in results if not r.error]
    lat = np.array([r.latency for r in ok]) if ok else np.zeros(1)
    steps = sum(r.time_steps for r in ok)
    return {"requests": len(results), "failed": len(results) - len(ok),
            "time_steps": int(steps), "makespan_s": round(makespan, 3),
            "time_steps_per_s": round(steps / makespan, 2) if makespan > 0 else 0.0,
            "latency_p50_s": round(float(np.percentile(lat, 50)), 3),
            "latency_p99_s": round(float(np.percentile(lat, 99)), 3)}


def poisson_requests(problems, rate, nrhs=32, seed=5, precision=api.PRECISION_MIXED):
    """Requests over the given problems with exponential inter-arrival times (rate per second;
    rate None: all at time 0), per-request load amplitudes U[0.5, 2] and rise times U[0, 10 dt]
    (half of them Heaviside)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for i, p in enumerate(problems):
        if rate is not None:
            t += rng.exponential(1.0 / rate)
        amp = rng.uniform(0.5, 2.0, nrhs)
        rise = np.where(rng.uniform(size=nrhs) < 0.5, 0.0, rng.uniform(0, 10 * p.dt, nrhs))
        reqs.append(Request(i, p, t, amp, rise, precision))
    return reqs
