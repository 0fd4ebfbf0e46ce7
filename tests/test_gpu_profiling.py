"""Instrumentation levels of tdbem_set_profiling: off, the history convolution only, every timed
kernel.  The level changes what is timed and counted, never the results."""
import numpy as np
import pytest

import helpers
from tdbem import api, configs

pytestmark = pytest.mark.gpu
NRHS = 16


def _solve(p, level):
    amp, rise = configs.rhs_set(NRHS, p.dt, seed=3)
    with api.Context.from_problem(p) as ctx:
        ctx.set_precision(api.PRECISION_MIXED)
        if level:
            ctx.set_profiling(True, conv_only=level == api.PROFILE_CONV)
        ctx.build_lag_store()
        ctx.march_init(p.a1, p.a2, p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val, amp=amp,
                       t_rise=rise)
        ctx.march_run()
        return ctx.timings(), [ctx.history(r) for r in range(NRHS)]


def test_profiling_levels(gpu):
    p = helpers.small_bar(numstep=24)
    print("[gpu case] profiling levels, bar h=0.5, 24 steps", flush=True)
    t0, h0 = _solve(p, 0)
    tc, hc = _solve(p, api.PROFILE_CONV)
    ta, ha = _solve(p, api.PROFILE_ALL)
    steps_with_conv = p.numstep - 2   # step 1 (the n == 2 solve) has no history sum
    assert t0["conv_launches"] == 0 and t0["lag_launches"] == 0 and t0["conv_ms"] == 0
    assert tc["conv_launches"] == steps_with_conv and tc["conv_ms"] > 0
    assert tc["solve_ms"] == 0 and tc["reduce_ms"] == 0
    assert tc["lag_launches"] == 0 and tc["active_points"] == 0
    assert ta["conv_launches"] == steps_with_conv and ta["conv_ms"] > 0
    assert ta["solve_ms"] > 0 and ta["reduce_ms"] > 0
    assert ta["lag_launches"] > 0 and ta["lag_kernel_ms"] > 0 and ta["active_points"] > 0
    for t in (t0, tc, ta):
        assert t["store_ms"] > 0 and t["init_ms"] > 0 and t["march_ms"] > 0
    for a, b, c in zip(h0, hc, ha):
        for k in range(2):
            assert np.array_equal(a[k], b[k]) and np.array_equal(a[k], c[k])
