"""Invariants and closed forms that pin the oracle independently of the reference's code:
shape-function identities, the static Kelvin limit of the Stokes kernel, a brute-force time
quadrature of the Heaviside-load Stokes displacement, and the zero/static windows that the GPU
shell store relies on."""
import numpy as np
import pytest
from scipy import integrate

import oracle_lib as O

E, NU, RHO = 1.3, 0.27, 1.1


def speeds():
    m = O.material(E, NU, RHO)
    return m.cp, m.cs, m.mu


@pytest.mark.parametrize("etype", [1, 2])
def test_shape_partition_of_unity_and_planar_area(etype):
    X4 = np.array([[0, 0, 0], [2, 0, 0], [2.5, 1.5, 0], [0, 1, 0]], float)
    X = X4 if etype == 1 else np.vstack([X4, 0.5 * (X4 + np.roll(X4, -1, axis=0))])
    gx = np.array([0.6612093864662645, -0.6612093864662645, -0.2386191860831969,
                   0.2386191860831969, -0.9324695142031521, 0.9324695142031521])
    gw = np.array([0.3607615730481386, 0.3607615730481386, 0.4679139345726910,
                   0.4679139345726910, 0.1713244923791704, 0.1713244923791704])
    area = 0.0
    for a, wa in zip(gx, gw):
        for b, wb in zip(gx, gw):
            N, x, n, J = O.shape_eval(etype, X, a, b)
            assert abs(N.sum() - 1) < 1e-14
            np.testing.assert_allclose(x, N @ X, atol=1e-14)     # linear field reproduction
            np.testing.assert_allclose(n, [0, 0, 1], atol=1e-14)
            area += wa * wb * J
    # shoelace area of the planar quad
    P = X4[:, :2]
    exact = 0.5 * abs(np.dot(P[:, 0], np.roll(P[:, 1], -1)) - np.dot(P[:, 1], np.roll(P[:, 0], -1)))
    assert abs(area - exact) < 1e-13 * exact


def kelvin_u(d, mu, nu):
    r = np.linalg.norm(d)
    g = d / r
    return ((3 - 4 * nu) * np.eye(3) + np.outer(g, g)) / (16 * np.pi * mu * (1 - nu) * r)


def test_heaviside_kernel_static_limit_is_kelvin():
    """Once both waves have passed (t - r/c_s >= dt) the hat-weighted Heaviside response is the
    static Kelvin solution times the hat integral dt/2, for displacement and traction."""
    cp, cs, mu = speeds()
    rng = np.random.default_rng(1)
    for _ in range(20):
        d = rng.normal(size=3)
        n = rng.normal(size=3)
        r = np.linalg.norm(d)
        dt = rng.uniform(0.01, 0.3)
        t = r / cs + dt * rng.uniform(1.0, 5.0)
        DU1, DU2, DT1, DT2, K1, _ = O.green(2, E, NU, RHO, d, n, t, dt)
        UK = kelvin_u(d, mu, NU) * dt / 2
        np.testing.assert_allclose(DU1, UK, rtol=1e-10, atol=1e-12 * np.abs(UK).max())
        np.testing.assert_allclose(DU2, UK, rtol=1e-10, atol=1e-12 * np.abs(UK).max())
        # Kelvin_Linear is the traction times dt/2
        np.testing.assert_allclose(DT1, K1, rtol=1e-9, atol=1e-11 * np.abs(K1).max())
        np.testing.assert_allclose(DT2, K1, rtol=1e-9, atol=1e-11 * np.abs(K1).max())


def stokes_heaviside_u(d, tp, cp, cs, rho):
    """Displacement of a Heaviside point force (Stokes' solution, e.g. Aki & Richards eq. 4.23
    with X0 = H): near field integral of tau over [r/cp, min(t, r/cs)] plus the two far fields."""
    r = np.linalg.norm(d)
    g = d / r
    if tp <= r / cp:
        return np.zeros((3, 3))
    nfield = 0.5 * (min(tp, r / cs) ** 2 - (r / cp) ** 2)
    u = (3 * np.outer(g, g) - np.eye(3)) * nfield / r ** 3
    u += np.outer(g, g) / (cp ** 2 * r)
    if tp > r / cs:
        u -= (np.outer(g, g) - np.eye(3)) / (cs ** 2 * r)
    return u / (4 * np.pi * rho)


def test_heaviside_displacement_matches_brute_force_time_quadrature():
    cp, cs, mu = speeds()
    rng = np.random.default_rng(2)
    for _ in range(12):
        d = rng.normal(size=3)
        r = np.linalg.norm(d)
        dt = rng.uniform(0.05, 0.4)
        t = r / cp + rng.uniform(-0.5, 4.0) * dt
        DU1, DU2, *_ = O.green(2, E, NU, RHO, d, [0, 0, 1], t, dt)
        for k, (phi, DU) in enumerate(((lambda s: (dt - s) / dt, DU1), (lambda s: s / dt, DU2))):
            ref = np.zeros((3, 3))
            brk = [x for x in (t - r / cp, t - r / cs) if 0 < x < dt]
            for i in range(3):
                for j in range(3):
                    ref[i, j] = integrate.quad(
                        lambda s: phi(s) * stokes_heaviside_u(d, t - s, cp, cs, RHO)[i, j], 0, dt,
                        points=brk or None, epsabs=0, epsrel=1e-13, limit=200)[0]
            scale = max(np.abs(ref).max(), 1e-30)
            assert np.abs(DU - ref).max() <= 1e-10 * scale, (k, t, np.abs(DU - ref).max() / scale)


def test_step_kernel_windows():
    """Step kernel: exactly zero until the P wave arrives; zero up to rounding once both waves are
    static; equal to HL(t) - HL(t - dt) in between."""
    cp, cs, mu = speeds()
    d = np.array([0.4, -0.7, 0.5])
    n = np.array([0.1, 0.2, 1.0])
    r = np.linalg.norm(d)
    dt = 0.05
    for t in np.arange(dt, r / cp + 1e-12, dt):
        for M in O.green(4, E, NU, RHO, d, n, t, dt)[:4]:
            assert np.all(M == 0.0)
    big = np.abs(O.green(2, E, NU, RHO, d, n, r / cs + 3 * dt, dt)[0]).max()
    for t in np.arange(r / cs + 2 * dt + 1e-9, r / cs + 20 * dt, dt):
        for M in O.green(4, E, NU, RHO, d, n, t, dt)[:4]:
            assert np.abs(M).max() <= 1e-12 * big * (t / dt) ** 2
    t = r / cp + 1.5 * dt
    a = O.green(4, E, NU, RHO, d, n, t, dt)
    b = O.green(2, E, NU, RHO, d, n, t, dt)
    c = O.green(2, E, NU, RHO, d, n, t - dt, dt)
    for q in range(4):
        np.testing.assert_allclose(a[q], b[q] - c[q], rtol=1e-12, atol=1e-15)


def test_bspline_kelvin_weights_integrate_the_spline():
    """Int_CubicB_Linear weights times the Kelvin traction: for t >= dt the width-0.7dt spline is
    fully inside the hat window in the past, so g1 + g2 = integral of the spline = 0.7 dt * (1/4
    of its peak-normalised area, 1/6 * 4 ... ) is fixed; check against direct quadrature."""
    d = np.array([0.3, 0.2, -0.9])
    n = np.array([0.0, 0.0, 1.0])
    dt = 0.2
    bw = 0.7 * dt

    def spline(x):
        x = x / bw
        if x < 0 or x >= 1:
            return 0.0
        if x < 0.25:
            return 32 / 3 * x ** 3
        if x < 0.5:
            return -32 * x ** 3 + 32 * x ** 2 - 8 * x + 2 / 3
        if x < 0.75:
            return 32 * x ** 3 - 64 * x ** 2 + 40 * x - 22 / 3
        return -32 / 3 * x ** 3 + 32 * x ** 2 - 32 * x + 32 / 3

    for t in (0.5 * dt, dt, 1.3 * dt):
        K1, K2 = O.green(5, E, NU, RHO, d, n, t, dt)[4:6]
        KT = np.zeros(9)
        # KelvinTraction = Kelvin_Linear * 2 / dt
        KT = O.green(4, E, NU, RHO, d, n, t, dt)[4] * 2 / dt
        g1 = integrate.quad(lambda s: (dt - s) / dt * spline(t - s), 0, dt, limit=200,
                            points=[t - bw * f for f in (0, .25, .5, .75, 1) if 0 < t - bw * f < dt])[0]
        # the reference integrates with 20 Gauss points; compare to the exact integral loosely
        assert np.abs(K1 - KT * g1).max() <= 2e-3 * np.abs(KT).max() * max(abs(g1), 1e-3)


def test_ramp_kernel_zero_before_arrival_and_static_after():
    """Ramped-Heaviside load of slope 2/dt: nothing before the P wave arrives; once the S wave
    has passed the end of the ramp (t - dt/2 >= r/c_s) the response is the static Kelvin
    displacement and traction."""
    cp, cs, mu = speeds()
    rng = np.random.default_rng(5)
    for _ in range(20):
        d = rng.normal(size=3)
        n = rng.normal(size=3)
        r = np.linalg.norm(d)
        dt = rng.uniform(0.01, 0.3)
        U, T, g = O.ramp_point(E, NU, RHO, d, n, r / cp * rng.uniform(0.2, 0.999), dt)
        assert np.all(U == 0) and np.all(T == 0)
        t = r / cs + dt / 2 + dt * rng.uniform(0.0, 6.0)
        U, T, g = O.ramp_point(E, NU, RHO, d, n, t, dt)
        UK = kelvin_u(d, mu, NU)
        TK = O.green(4, E, NU, RHO, d, n, t, dt)[4] * 2 / dt      # Kelvin_Linear = T^K dt/2
        np.testing.assert_allclose(U, UK, rtol=1e-9, atol=1e-11 * np.abs(UK).max())
        np.testing.assert_allclose(T, TK, rtol=1e-8, atol=1e-10 * np.abs(TK).max())


def test_ramp_displacement_is_the_ramp_convolution_of_the_heaviside_response():
    """U_ramp(t) = k * integral over [0, 1/k] of U_Heaviside(t - s) ds (k = 2/dt): Duhamel's
    principle for the load g(t) = min(max(k t, 0), 1), by brute-force quadrature of the
    textbook Heaviside-load displacement."""
    cp, cs, mu = speeds()
    rng = np.random.default_rng(6)
    for _ in range(10):
        d = rng.normal(size=3)
        r = np.linalg.norm(d)
        dt = rng.uniform(0.05, 0.4)
        k = 2 / dt
        t = r / cp + rng.uniform(-0.2, 4.0) * dt
        U, _, _ = O.ramp_point(E, NU, RHO, d, [0, 0, 1], t, dt)
        brk = [x for x in (t - r / cp, t - r / cs) if 0 < x < 1 / k]
        ref = np.zeros((3, 3))
        for i in range(3):
            for j in range(3):
                ref[i, j] = k * integrate.quad(
                    lambda s: stokes_heaviside_u(d, t - s, cp, cs, RHO)[i, j], 0, 1 / k,
                    points=brk or None, epsabs=0, epsrel=1e-13, limit=200)[0]
        scale = max(np.abs(ref).max(), 1e-30)
        assert np.abs(U - ref).max() <= 1e-10 * scale, (t, np.abs(U - ref).max() / scale)


def test_ramp_kelvin_weights_are_the_hat_integrals_of_the_ramp():
    """Int_RampedHeaviside_Linear: g1, g2 = integral over the step of phi_1, phi_2 times the
    ramp load at t - tau (40-point rule), against adaptive quadrature.  The rule is exact for the
    piecewise-polynomial integrand once no kink of the load lies inside the step, and within a
    few 1e-5 dt when one does (the reference's own quadrature error there)."""
    for dt in (0.1, 0.37):
        k = 2 / dt
        for t in np.arange(0.25, 4.0, 0.25) * dt:
            _, _, g = O.ramp_point(E, NU, RHO, [1, 0, 0], [0, 0, 1], t, dt)
            load = lambda s: min(max(k * (t - s), 0.0), 1.0)
            kinks = [x for x in (t, t - 1 / k) if 0 < x < dt]
            e1 = integrate.quad(lambda s: (dt - s) / dt * load(s), 0, dt, points=kinks or None)[0]
            e2 = integrate.quad(lambda s: s / dt * load(s), 0, dt, points=kinks or None)[0]
            tol = 1e-4 * dt if kinks else 1e-13 * dt
            assert abs(g[0] - e1) <= tol and abs(g[1] - e2) <= tol, (t / dt, g, e1, e2)
