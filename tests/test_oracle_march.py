"""Pins the oracle's time marching (src/SolverStepLinearT.cpp:314-544).  PETSc is not installed,
so the reference solver cannot run anywhere here; its recurrence is pinned instead by:
  - the fused one-pass form derived independently (or_march_fused) agreeing with the literal
    restatement of the three averaged lag loops;
  - alpha1 = alpha2 = 0 reducing to plain collocation, checked against a brute-force solve of the
    whole block-lower-triangular space-time system (the structure of Solve_Direct, :185-308);
  - linearity in the load and the alpha range check."""
import numpy as np
import pytest

import helpers
import oracle_lib as O


@pytest.fixture(scope="module")
def bar_lags():
    p = helpers.small_bar(numstep=16)
    K = O.max_step(p.E, p.nu, p.rho, p.coords, p.dt)
    MG, MH = O.build_lags(4, 1, 6, p.E, p.nu, p.rho, p.coords, p.ien, 0, p.dt, p.numstep, K)
    return p, MG, MH


def test_build_lags_follows_update_gh_linear(bar_lags):
    p, MG, MH = bar_lags
    steps = [O.assemble_step(4, 1, 6, p.E, p.nu, p.rho, p.coords, p.ien, 0, 0, p.nnd - 1,
                             n * p.dt, p.dt, n) for n in (1, 2, 3)]
    # M_0 = phi_2(step 1); M_1 = phi_1(step 1) + phi_2(step 2); M_2 = phi_1(2) + phi_2(3)
    np.testing.assert_array_equal(MG[0], steps[0][1])
    np.testing.assert_array_equal(MH[1], steps[0][2] + steps[1][3])
    np.testing.assert_array_equal(MG[2], steps[1][0] + steps[2][1])


@pytest.mark.parametrize("a1,a2", [(0.25, 0.25), (0.0, 0.0), (0.4, 0.1), (0.1, 0.6)])
def test_literal_and_fused_marches_agree(bar_lags, a1, a2):
    p, MG, MH = bar_lags
    U1, T1 = O.march(MG, MH, p.numstep, p.dt, p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val, a1, a2)
    U2, T2 = O.march(MG, MH, p.numstep, p.dt, p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val, a1, a2,
                     fused=True)
    assert helpers.rel_fro(U2, U1) < 1e-12
    assert helpers.rel_fro(T2, T1) < 1e-12


def test_collocation_equals_space_time_solve(bar_lags):
    """a1 = a2 = 0: each step solves sum_{k} (M^H_k u_{s-k} - M^G_k t_{s-k}) = 0 with the known
    parts moved right.  Here all steps are solved at once as one block-triangular system."""
    p, MG, MH = bar_lags
    S = 10
    N = 3 * p.nnd
    K = MG.shape[0] - 1
    ubc = np.zeros(N, bool)
    ubc[p.ubc_dof] = True
    uval = np.zeros(N)
    uval[p.ubc_dof] = p.ubc_val
    tval = np.zeros(N)
    tval[p.fbc_dof] = p.fbc_val
    tval[ubc] = 0.0
    # unknown vector per step: x_s (u at non-UBC dofs, t at UBC dofs); 1-based steps 1..S
    A = np.zeros((S * N, S * N))
    b = np.zeros(S * N)
    for s in range(1, S + 1):
        rows = slice((s - 1) * N, s * N)
        for k in range(0, min(s - 1, K) + 1):
            m = s - k  # history index
            # lag-0 matrix is M_0 at the first solve and the same M_0 (a1 = a2 = 0) afterwards
            Hk, Gk = MH[k], MG[k]
            cols = slice((m - 1) * N, m * N)
            # H u - G t: unknown u at non-UBC columns, unknown t at UBC columns
            A[rows, cols] += np.where(ubc[None, :], -Gk, Hk)
            # known parts: u = uval at UBC, t = tval elsewhere (traction free where not FBC)
            b[rows] -= Hk @ np.where(ubc, uval, 0.0) - Gk @ np.where(ubc, 0.0, tval)
    x = np.linalg.solve(A, b).reshape(S, N)
    U, T = O.march(MG, MH, S + 1, p.dt, p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val, 0.0, 0.0)
    u_bf = np.where(ubc[None, :], uval[None, :], x)
    t_bf = np.where(ubc[None, :], x, tval[None, :])
    t_bf[:, N - 1] = np.where(ubc[N - 1], 0.0, t_bf[:, N - 1])
    assert helpers.rel_fro(U[1:S + 1], u_bf) < 1e-10
    assert helpers.rel_fro(T[1:S + 1], t_bf) < 1e-10


def test_march_is_linear_in_the_load(bar_lags):
    p, MG, MH = bar_lags
    U1, T1 = O.march(MG, MH, p.numstep, p.dt, p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val, amp=1.0)
    U2, T2 = O.march(MG, MH, p.numstep, p.dt, p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val, amp=2.5)
    assert helpers.rel_fro(U2, 2.5 * U1) < 1e-13
    assert helpers.rel_fro(T2, 2.5 * T1) < 1e-13


def test_alpha_out_of_range_is_rejected(bar_lags):
    p, MG, MH = bar_lags
    with pytest.raises(ValueError, match="Alpha is out of range"):
        O.march(MG, MH, 5, p.dt, p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val, 0.7, 0.5)
