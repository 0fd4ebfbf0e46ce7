"""Mixed-precision march (split-BF16 MFMA convolution): lane-layout self-test of the MFMA
builtin, then agreement with the plain FP64 path per right-hand side at the stated tolerance
(relative Frobenius error of the u and t histories <= 1e-6)."""
import numpy as np
import pytest

import helpers
from tdbem import api

pytestmark = pytest.mark.gpu
MIXED_TOL = 1e-6


def test_mfma_lane_layout(gpu):
    rng = np.random.default_rng(0)
    A = rng.integers(-8, 9, size=(16, 32)).astype(np.float32)
    B = rng.integers(-8, 9, size=(32, 16)).astype(np.float32)
    print("[gpu case] mfma self-test", flush=True)
    D = api.mfma_selftest(A, B)
    np.testing.assert_array_equal(D, A @ B)


def test_mfma_f64_lane_layout(gpu):
    rng = np.random.default_rng(1)
    A = rng.standard_normal((16, 4))
    B = rng.standard_normal((4, 16))
    print("[gpu case] mfma f64 self-test", flush=True)
    D = api.mfma_f64_selftest(A, B)
    np.testing.assert_allclose(D, A @ B, rtol=1e-14, atol=1e-14)


def _fp64_histories(p, amp, rise):
    out = []
    for r0 in range(0, len(amp), 8):
        with api.Context.from_problem(p) as ctx:
            ctx.march_init(p.a1, p.a2, p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val,
                           amp=amp[r0:r0 + 8], t_rise=rise[r0:r0 + 8])
            ctx.march_run()
            out += [ctx.history(r) for r in range(8)]
    return out


@pytest.mark.parametrize("case,nrhs", [("sphere", 16), ("bar", 32)])
def test_mixed_march_matches_fp64(gpu, case, nrhs):
    p = {"sphere": helpers.small_sphere, "bar": helpers.small_bar}[case](numstep=40)
    rng = np.random.default_rng(3)
    amp = rng.uniform(0.5, 2.0, nrhs)
    rise = np.where(rng.uniform(size=nrhs) < 0.5, 0.0, rng.uniform(0, 10 * p.dt, nrhs))
    print(f"[gpu case] mixed march {case} nrhs={nrhs}", flush=True)
    with api.Context.from_problem(p) as ctx:
        ctx.set_precision(api.PRECISION_MIXED)
        ctx.march_init(p.a1, p.a2, p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val, amp=amp,
                       t_rise=rise)
        ctx.march_run()
        mixed = [ctx.history(r) for r in range(nrhs)]
    ref = _fp64_histories(p, amp, rise)
    errs = [max(helpers.rel_fro(m[0], r[0]), helpers.rel_fro(m[1], r[1]))
            for m, r in zip(mixed, ref)]
    print(f"mixed vs fp64 max rel err {max(errs):.3e}")
    assert max(errs) <= MIXED_TOL


@pytest.mark.parametrize("case", ["sphere", "bar"])
def test_mixed_tile_store_is_the_fp64_store_rounded(gpu, case):
    """The mixed-precision build assembles in FP64 and keeps every lag only as FP32 sub-tiles
    (filled straight from the lag kernel's slots, no FP64 copy of the whole store): each stored
    value is the FP64 store's value rounded once to FP32, so the lags agree to FP32 rounding
    (relative Frobenius over all lags <= 1e-7, and not exactly); switching back to FP64
    rebuilds the FP64 store exactly."""
    p = {"sphere": helpers.small_sphere, "bar": helpers.small_bar}[case](numstep=40)
    print(f"[gpu case] mixed tile store {case}", flush=True)
    with api.Context.from_problem(p) as c64, api.Context.from_problem(p) as c32:
        c64.build_lag_store()
        c32.set_precision(api.PRECISION_MIXED)
        c32.build_lag_store()
        num = den = 0.0
        for k in range(c64.max_step + 1):
            G64, H64 = c64.get_lag(k)
            G32, H32 = c32.get_lag(k)
            # the same FP64 sums, rounded once
            assert np.array_equal(G32, G64.astype(np.float32).astype(np.float64)), k
            assert np.array_equal(H32, H64.astype(np.float32).astype(np.float64)), k
            num += np.linalg.norm(G32 - G64) ** 2 + np.linalg.norm(H32 - H64) ** 2
            den += np.linalg.norm(G64) ** 2 + np.linalg.norm(H64) ** 2
        err = np.sqrt(num / den)
        print(f"tile store rel err {err:.3e}")
        assert 0 < err <= 1e-7
        c32.set_precision(api.PRECISION_FP64)
        c32.build_lag_store()
        for k in range(c64.max_step + 1):
            G64, H64 = c64.get_lag(k)
            G, H = c32.get_lag(k)
            assert np.array_equal(G, G64) and np.array_equal(H, H64)


def test_mixed_store_rebuilt_on_one_context(gpu):
    """Repeated solves on one context (as the benchmark times them): the second and third builds
    of the mixed-precision store reuse or redo the tile plan and give the same histories bit for
    bit (regression: a rebuild once launched the tile gather without its sub-tile list)."""
    p = helpers.small_sphere(numstep=20)
    rng = np.random.default_rng(4)
    amp = rng.uniform(0.5, 2.0, 16)
    rise = np.zeros(16)
    print("[gpu case] mixed store rebuilt three times", flush=True)
    with api.Context.from_problem(p) as ctx:
        ctx.set_precision(api.PRECISION_MIXED)
        runs = []
        for _ in range(3):
            ctx.build_lag_store()
            ctx.march_init(p.a1, p.a2, p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val, amp=amp,
                           t_rise=rise)
            ctx.march_run()
            runs.append([ctx.history(r) for r in (0, 15)])
    for run in runs[1:]:
        for (u, t), (u0, t0) in zip(run, runs[0]):
            assert np.array_equal(u, u0) and np.array_equal(t, t0)
