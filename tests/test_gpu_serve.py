"""Continuous-batching engine (tdbem.serve.Engine): every request's histories are the same alone
and in the mix (bit for bit: a scenario's convolution is cut by its own work only), the mix's
solves share the batch's launches, admission respects the HBM budget, and the byte predictions
that admission uses match what the device reports."""
import numpy as np
import pytest

import helpers
from tdbem import api, configs, serve

pytestmark = pytest.mark.gpu
NRHS = 16


def _problems():
    hs = configs.halfspace_scenarios(count=1, seed=7, el_range=(120, 160), step_range=(24, 24))[0]
    return [helpers.small_bar(numstep=24), helpers.small_sphere(numstep=28), hs,
            helpers.small_bar(numstep=20, h=1.0)]


def test_results_same_alone_and_in_the_mix(gpu):
    probs = _problems()
    print("[gpu case] engine: alone vs mix", flush=True)
    reqs = serve.poisson_requests(probs, rate=None, nrhs=NRHS, seed=5)
    alone = []
    for r in reqs:
        res, summary = serve.Engine(nrhs=NRHS, setup_workers=1, keep_histories=True).run([r])
        assert not res[0].error, res[0].error
        assert summary["engine_max_members"] == 1
        alone.append(res[0].histories)
    mixed, summary = serve.Engine(nrhs=NRHS, setup_workers=4, steps_per_graph=4,
                                  keep_histories=True).run(reqs)
    assert summary["failed"] == 0
    for a, m in zip(alone, mixed):
        assert len(m.histories) == NRHS
        for (ua, ta), (um, tm) in zip(a, m.histories):
            assert np.array_equal(um, ua) and np.array_equal(tm, ta)
    # the solves were stepped together: several members per launch, membership spans overlap
    assert summary["engine_max_members"] >= 2
    spans = sorted((m.joined, m.finished) for m in mixed)
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))


def test_engine_matches_the_single_context_march(gpu):
    """A request's histories from the engine equal the same load cases marched on their own
    context (march_run), bit for bit."""
    p = helpers.small_sphere(numstep=30)
    r = serve.poisson_requests([p], rate=None, nrhs=NRHS, seed=9)[0]
    print("[gpu case] engine vs single context", flush=True)
    res, _ = serve.Engine(nrhs=NRHS, keep_histories=True).run([r])
    with api.Context.from_problem(p) as ctx:
        ctx.set_precision(api.PRECISION_MIXED)
        ctx.march_init(p.a1, p.a2, p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val, amp=r.amp,
                       t_rise=r.rise)
        ctx.march_run()
        for k in range(NRHS):
            U, T = ctx.history(k)
            assert np.array_equal(res[0].histories[k][0], U)
            assert np.array_equal(res[0].histories[k][1], T)


def test_admission_respects_the_hbm_budget(gpu):
    probs = _problems()[:3]
    reqs = serve.poisson_requests(probs, rate=None, nrhs=NRHS, seed=6)
    print("[gpu case] engine admission budget", flush=True)
    # a budget smaller than any two requests together: they must run one at a time
    res, summary = serve.Engine(nrhs=NRHS, hbm_budget=1.0, setup_workers=4).run(reqs)
    assert summary["failed"] == 0
    spans = sorted((r.admitted, r.finished) for r in res)
    for i in range(len(spans) - 1):
        assert spans[i + 1][0] >= spans[i][1]
    assert summary["engine_max_members"] == 1
    assert summary["latency_p99_s"] >= summary["latency_p50_s"] > 0
    # a budget that holds the two smallest requests' predicted peaks: more is never committed,
    # except by a request larger than the budget on its own (admitted only onto an idle device)
    peaks = sorted(r.predicted_bytes for r in res)
    budget = peaks[0] + peaks[1]
    res2, summary2 = serve.Engine(nrhs=NRHS, hbm_budget=budget, setup_workers=4).run(reqs)
    assert summary2["failed"] == 0
    assert summary2["engine_max_committed_bytes"] <= max(budget, peaks[-1])
    alone = [r for r in res2 if r.predicted_bytes > budget]
    for r in alone:
        others = [(o.admitted, o.finished) for o in res2 if o is not r]
        assert all(f <= r.admitted or a >= r.finished for a, f in others)


@pytest.mark.parametrize("case,precision,nrhs", [("sphere", api.PRECISION_MIXED, 32),
                                                 ("bar", api.PRECISION_MIXED, 16),
                                                 ("sphere", api.PRECISION_FP64, 8)])
def test_predicted_bytes_match_the_device(gpu, case, precision, nrhs):
    """tdbem_predict_bytes before anything is built, against hipMemGetInfo (pool trimmed) after
    the set-up: the resident prediction within 3% + 32 MiB, the peak not below it; everything is
    returned when the context is freed."""
    p = {"sphere": lambda: configs.sphere(n=12, numstep=200),
         "bar": lambda: configs.bar(h=0.25, numstep=100)}[case]()
    print(f"[gpu case] predicted bytes {case} precision={precision} nrhs={nrhs}", flush=True)
    with api.Context.from_problem(helpers.small_bar(numstep=8, h=1.0)) as warm:
        warm.march_init()   # loads the kernels' code before the measurement
    # and the BLAS library's (its kernels and workspace stay for the process, once per device)
    n = 1536
    api.dense_solve(np.eye(n) * 2.0 + 1e-3, np.ones((n, 1)))
    free0, _ = api.device_memory()
    ctx = api.Context.from_problem(p)
    ctx.set_precision(precision)
    resident, peak = ctx.predict_bytes(nrhs)
    amp, rise = configs.rhs_set(nrhs, p.dt, seed=2)
    ctx.build_lag_store()
    ctx.march_init(p.a1, p.a2, p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val, amp=amp, t_rise=rise)
    free1, _ = api.device_memory()
    used = free0 - free1
    print(f"{case}: predicted resident {resident / 2**20:.1f} MiB, peak {peak / 2**20:.1f} MiB,"
          f" device {used / 2**20:.1f} MiB")
    assert abs(used - resident) <= 0.03 * resident + (32 << 20)
    assert peak >= resident
    ctx.close()
    free2, _ = api.device_memory()
    assert free0 - free2 <= (16 << 20)
