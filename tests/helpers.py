"""Shared helpers for the tests: relative errors and small workloads."""
import numpy as np

from tdbem import configs, meshgen


def rel_fro(a, b):
    """||a - b||_F / ||b||_F (0 when both are zero)."""
    nb = np.linalg.norm(b)
    na = np.linalg.norm(a - b)
    if nb == 0:
        return float(na)
    return float(na / nb)


def rel_fro_set(got, ref):
    """Relative Frobenius error of a set of matrices taken together (e.g. all lags or all steps
    of one of G1/G2/H1/H2): sqrt(sum ||g - r||^2) / sqrt(sum ||r||^2).  Lags past the static
    time of every pair hold only rounding noise (~1e-14 of the peak lag); measured one by one they
    would compare noise with noise, so parity is stated over the whole set."""
    num = sum(float(np.linalg.norm(g - r)) ** 2 for g, r in zip(got, ref))
    den = sum(float(np.linalg.norm(r)) ** 2 for r in ref)
    return float(np.sqrt(num / den)) if den > 0 else float(np.sqrt(num))


def shell_parity(got, ref):
    """Parity of lag matrices that the GPU holds only inside each node pair's shell (exact zeros
    outside) with the oracle's dense ones, which -- like the reference -- evaluate the Step kernel
    HL(t) - HL(t - dt) at every lag: past a pair's static time (t - r/c_s >= dt at every
    quadrature point) the two terms are equal and the difference is rounding noise, eps t^2 / r^2
    at the near-singular points, growing with the number of lags.  Returns (in_shell, outside):
    the relative Frobenius error over the entries the GPU holds, and the norm of the oracle's
    entries where the GPU holds zeros, both relative to the oracle's norm over the whole set."""
    num = den = out = 0.0
    for g, r in zip(got, ref):
        zero = g == 0
        out += float(np.linalg.norm(r[zero])) ** 2
        num += float(np.linalg.norm((g - r)[~zero])) ** 2
        den += float(np.linalg.norm(r)) ** 2
    return float(np.sqrt(num / den)), float(np.sqrt(out / den))


def shell_masked(lags, got):
    """The oracle's lags with the entries where the GPU store holds zeros set to zero."""
    return np.stack([np.where(g == 0, 0.0, r) for g, r in zip(got, lags)])


def plane_1e():
    """geometry/HalfSpacePlane_1E.txt: one unit square element, E=5e7, nu=0.25, rho=1730,
    dt=5e-3, if_ext=1, traction 1 on the z dofs."""
    coords = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], dtype=np.float64)
    ien = np.array([[0, 1, 2, 3]], dtype=np.int32)
    return configs.Problem("plane1e", coords, ien, 5e7, 0.25, 1730.0, 5e-3, 100, 1,
                           np.zeros(0, np.int32), np.zeros(0),
                           np.array([2, 5, 8, 11], np.int32), np.ones(4))


def small_bar(numstep=40, h=0.5):
    return configs.bar(h=h, numstep=numstep)


def small_sphere(n=4, numstep=40):
    return configs.sphere(n=n, numstep=numstep)


def small_quadratic_bar(numstep=12):
    b = configs.bar(h=1.0, numstep=numstep)
    m = meshgen.to_serendipity(meshgen.Mesh(b.coords, b.ien))
    p = configs.Problem("bar8", m.coords, m.ien, b.E, b.nu, b.rho, b.dt, numstep, 0,
                        np.zeros(0, np.int32), np.zeros(0), np.zeros(0, np.int32), np.zeros(0),
                        etype=2)
    return p
