"""Batched march (tdbem_batch_*): scenarios of uneven size and step count stepped together by one
hipGraph of several steps.  Every graph replay, with eager work of another context and host
synchronisation between replays, is checked against each scenario marched on its own, and the
autotuner's choice leaves the results unchanged to 1e-10.  (The tunables change the order of FP64
sums; the FP32 history copy can turn such last-bit differences into FP32 roundings of single
values, 2^-24 of the value each, which the following steps carry on: 2e-12 has been seen over
these 18-37 steps, and ~3e-8 after 600 sphere steps, against the mixed path's 7.5e-8 distance
from FP64.)"""
import numpy as np
import pytest

import helpers
from tdbem import api, configs

pytestmark = pytest.mark.gpu
NRHS = 16
TOL = 1e-10


def _scenarios():
    hs = configs.halfspace_scenarios(count=2, seed=11, el_range=(110, 170), step_range=(22, 30))
    return [helpers.small_bar(numstep=30, h=0.5), helpers.small_sphere(numstep=37), hs[0],
            helpers.small_bar(numstep=18, h=1.0), hs[1]]


def _loads(p, seed):
    rng = np.random.default_rng(seed)
    amp = rng.uniform(0.5, 2.0, NRHS)
    rise = np.where(rng.uniform(size=NRHS) < 0.5, 0.0, rng.uniform(0, 10 * p.dt, NRHS))
    return amp, rise


def _start(p, seed):
    ctx = api.Context.from_problem(p)
    ctx.set_precision(api.PRECISION_MIXED)
    amp, rise = _loads(p, seed)
    ctx.march_init(p.a1, p.a2, p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val, amp=amp, t_rise=rise)
    return ctx


def _alone(probs):
    out = []
    for i, p in enumerate(probs):
        with _start(p, 100 + i) as ctx:
            ctx.march_run()
            out.append([ctx.history(r) for r in range(NRHS)])
    return out


def _check_prefix(ctx, ref, upto):
    for r in range(NRHS):
        U, T = ctx.history(r)
        Ur, Tr = ref[r]
        for a, b in ((U, Ur), (T, Tr)):
            err = np.linalg.norm(a[:upto + 1] - b[:upto + 1]) / max(np.linalg.norm(b), 1e-300)
            assert err <= TOL, (upto, r, err)


def test_batch_replays_match_scenarios_alone(gpu):
    probs = _scenarios()
    print("[gpu case] batch: scenarios alone", flush=True)
    ref = _alone(probs)
    ctxs = [_start(p, 100 + i) for i, p in enumerate(probs)]
    other = helpers.small_bar(numstep=12, h=1.0)
    print("[gpu case] batch: graph replays of 4 steps", flush=True)
    with api.Batch(ctxs) as b:
        replays = 0
        while True:
            n = b.run(steps_per_graph=4, max_replays=1)
            if n == 0:
                break
            replays += n
            # eager work of another context between replays
            with api.Context.from_problem(other) as oc:
                oc.march_init(other.a1, other.a2, other.ubc_dof, other.ubc_val, other.fbc_dof,
                              other.fbc_val)
                oc.march_run()
            for ctx, rf in zip(ctxs, ref):
                _check_prefix(ctx, rf, ctx.step_index)
        assert replays >= (max(p.numstep for p in probs) - 2) // 4
    for ctx, p in zip(ctxs, probs):
        assert ctx.step_index == p.numstep - 1
        ctx.close()


def test_batch_tuned_run_matches_and_finished_scenarios_stay(gpu):
    probs = _scenarios()
    ref = _alone(probs)
    ctxs = [_start(p, 100 + i) for i, p in enumerate(probs)]
    print("[gpu case] batch: tune then run", flush=True)
    with api.Batch(ctxs) as b:
        rep = b.tune()
        assert rep["rounds"] in (1, 2, 4, 8) and rep["solve_waves"] in (1024, 2048, 4096, 8192)
        assert rep["conv_ms"] > 0 and rep["solve_ms"] > 0
        b.run(steps_per_graph=16)
        first = [[c.history(r) for r in range(NRHS)] for c in ctxs]
        assert b.run(steps_per_graph=16) == 0      # everything finished: no launch
    for ctx, rf, p in zip(ctxs, ref, probs):
        assert ctx.step_index == p.numstep - 1
        _check_prefix(ctx, rf, p.numstep - 1)
    for ctx, f in zip(ctxs, first):
        for r in range(NRHS):
            U, T = ctx.history(r)
            assert np.array_equal(U, f[r][0]) and np.array_equal(T, f[r][1])
        ctx.close()


def test_batch_rejects_mixed_right_hand_side_counts(gpu):
    p = helpers.small_bar(numstep=10, h=1.0)
    a = _start(p, 1)
    c = api.Context.from_problem(p)
    c.set_precision(api.PRECISION_MIXED)
    amp, rise = configs.rhs_set(32, p.dt, seed=2)
    c.march_init(p.a1, p.a2, p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val, amp=amp, t_rise=rise)
    with pytest.raises(api.TdbemError):
        api.Batch([a, c])
    a.close()
    c.close()


def test_concurrent_set_up_matches_sequential(gpu):
    """api.prepare_contexts runs the set-ups of several scenarios on host threads at once (each
    context on its own stream); the marches give the same histories as set-ups done one by
    one."""
    probs = _scenarios()
    print("[gpu case] concurrent set-up", flush=True)
    ref = _alone(probs)
    loads = [_loads(p, 100 + i) for i, p in enumerate(probs)]
    ctxs = api.prepare_contexts(probs, loads, precision=api.PRECISION_MIXED, workers=4)
    try:
        for ctx, r in zip(ctxs, ref):
            ctx.march_run()
            for k in range(NRHS):
                U, T = ctx.history(k)
                np.testing.assert_array_equal(U, r[k][0])
                np.testing.assert_array_equal(T, r[k][1])
    finally:
        for c in ctxs:
            c.close()


def test_batch_members_joining_at_either_step_parity(gpu):
    """The mixed convolution is time-blocked: at step s it convolves the far lags of parity
    s mod 2 for steps s and s + 1 and carries the step-(s + 1) part to the next step.  Members
    that marched alone for an even or an odd number of steps before joining (so the batch holds
    both parities at once), and that leave and march alone again, reproduce their march alone,
    bit for bit at the single-context tunables (rounds 4, 4096 solve wavefronts)."""
    probs = _scenarios()[:4]
    ref = _alone(probs)
    ctxs = [_start(p, 100 + i) for i, p in enumerate(probs)]
    print("[gpu case] batch: members joining at steps of both parities", flush=True)
    for i, ctx in enumerate(ctxs):
        ctx.march_run(3 + i)            # steps 1 .. 3 + i alone: parities 1, 0, 1, 0
    with api.Batch(ctxs[:2]) as b:
        b.set_params(4, 4096)
        b.run(steps_per_graph=5, max_replays=1)
        b.add(ctxs[2])
        b.add(ctxs[3])
        b.run(steps_per_graph=3, max_replays=2)
        b.remove(ctxs[0])
        b.run(steps_per_graph=4, max_replays=1)
    for ctx, rf, p in zip(ctxs, ref, probs):
        ctx.march_run()                 # the rest alone
        assert ctx.step_index == p.numstep - 1
        for r in range(NRHS):
            U, T = ctx.history(r)
            np.testing.assert_array_equal(U, rf[r][0])
            np.testing.assert_array_equal(T, rf[r][1])
        ctx.close()
