"""Serving loop: each request's result is the same alone and in the concurrent mix, admission
respects the HBM budget, and latencies/throughput are reported."""
import numpy as np
import pytest

import helpers
from tdbem import api, configs, serve

pytestmark = pytest.mark.gpu


def _problems():
    hs = configs.halfspace_scenarios(count=1, seed=7, el_range=(120, 160), step_range=(24, 24))[0]
    return [helpers.small_bar(numstep=24), helpers.small_sphere(numstep=28), hs,
            helpers.small_bar(numstep=20, h=1.0)]


def test_results_same_alone_and_in_the_mix(gpu):
    probs = _problems()
    print("[gpu case] serve alone vs mix", flush=True)
    reqs = serve.poisson_requests(probs, rate=1e6, nrhs=16, seed=5)
    for r in reqs:
        r.arrival = 0.0
    alone = []
    for r in reqs:
        res, _ = serve.Server(max_active=1, keep_histories=True).run([r])
        assert not res[0].error, res[0].error
        alone.append(res[0].histories)
    mixed, summary = serve.Server(max_active=4, keep_histories=True).run(reqs)
    assert summary["failed"] == 0
    for a, m in zip(alone, mixed):
        for (ua, ta), (um, tm) in zip(a, m.histories):
            assert helpers.rel_fro(um, ua) <= 1e-12
            assert helpers.rel_fro(tm, ta) <= 1e-12
    # the mix overlapped: some request was admitted before another finished
    spans = sorted((m.admitted, m.finished) for m in mixed)
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))


def test_admission_respects_the_hbm_budget(gpu):
    probs = _problems()[:3]
    reqs = serve.poisson_requests(probs, rate=1e6, nrhs=16, seed=6)
    for r in reqs:
        r.arrival = 0.0
    print("[gpu case] serve admission budget", flush=True)
    # a budget smaller than any two requests together: they must run one at a time
    res, summary = serve.Server(hbm_budget=1.0, max_active=4).run(reqs)
    assert summary["failed"] == 0
    spans = sorted((r.admitted, r.finished) for r in res)
    for i in range(len(spans) - 1):
        assert spans[i + 1][0] >= spans[i][1] - 1e-3
    assert summary["latency_p99_s"] >= summary["latency_p50_s"] > 0
