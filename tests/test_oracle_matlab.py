"""Pins the oracle's averaged march (or_march_literal, the restatement of
SolverStepLinearT::Solve_Averaging) against the reference's own independent restatement of the
same scheme in Matlab: the lag database built step by step as FE-BE Matlab/FE_BE_driver.m:88-131
builds it from the per-step G1/G2 files, and the system matrix and history sums of
FE-BE Matlab/BEM_averaging.m:19-105, whose coefficients are 4x those of a1 = a2 = 1/4 (so both
sides of every step's system are scaled by 4 and the solutions agree).

Matlab/Octave is absent, so the two .m functions are transcribed below with their 1-based
indices kept (cell G_DB{m} -> G_DB[m], row Dis_DB(m,:) -> x_{m-1}).  The boundary handling
(SetLoad, ExchangeColumns, the post-processing of SolverStepLinearT.cpp:512-536) is the part both
share and is written here from the C++ (the Matlab driver couples to an FE region instead)."""
import numpy as np

import helpers
import oracle_lib as O


def _per_step_blocks(p, last):
    """phi_1 and phi_2 parts of steps 1..last (G_H_Driver with t = n dt, main.cpp:89-105)."""
    out = {}
    for n in range(1, last + 1):
        G1, G2, H1, H2 = O.assemble_step(4, p.etype, p.nq, p.E, p.nu, p.rho, p.coords, p.ien,
                                         p.if_ext, 0, p.nnd - 1, n * p.dt, p.dt, n)
        out[n] = (G1, G2, H1, H2)
    return out


def _matlab_database(blocks, num_matrix):
    """FE_BE_driver.m:88-131.  The driver's G1 files hold the phi_2 part and its G2 files the
    phi_1 part of each step (WriteGH's naming, main.cpp:113-116).  Returns G_DB, H_DB indexed
    1..num_matrix+1 as the cells are after the last update."""
    G1_DB = {n: blocks[n][1] for n in blocks}
    G2_DB = {n: blocks[n][0] for n in blocks}
    H1_DB = {n: blocks[n][3] for n in blocks}
    H2_DB = {n: blocks[n][2] for n in blocks}
    G_DB, H_DB = {}, {}
    G_DB[1] = G1_DB[1].copy()
    G_DB[2] = G2_DB[1].copy()
    H_DB[1] = H1_DB[1].copy()
    H_DB[2] = H2_DB[1].copy()
    for curr_step in range(2, num_matrix + 1):
        G_DB[curr_step] = G_DB[curr_step] + G1_DB[curr_step]
        G_DB[curr_step + 1] = G2_DB[curr_step].copy()
        H_DB[curr_step] = H_DB[curr_step] + H1_DB[curr_step]
        H_DB[curr_step + 1] = H2_DB[curr_step].copy()
    return G_DB, H_DB


def _bem_averaging(G_DB, H_DB, Dis_DB, Trac_DB, curr_step, max_step):
    """BEM_averaging.m:19-105 (the tail terms multiply Dis_DB(1,:) = Trac_DB(1,:) = 0 and are
    left out)."""
    if curr_step == 2:
        return G_DB[1], H_DB[1], np.zeros(G_DB[1].shape[0])
    G_curr = 4 * G_DB[1] + G_DB[2]
    H_curr = 4 * H_DB[1] + H_DB[2]
    rhs = np.zeros(G_curr.shape[0])
    for f_i in range(3, min(max_step, curr_step + 1) + 1):          # n+1 step
        rhs += G_DB[f_i] @ Trac_DB[curr_step + 2 - f_i - 1]
        rhs -= H_DB[f_i] @ Dis_DB[curr_step + 2 - f_i - 1]
    for f_i in range(2, min(max_step, curr_step - 1) + 1):          # n step
        rhs += 2 * G_DB[f_i] @ Trac_DB[curr_step + 1 - f_i - 1]
        rhs -= 2 * H_DB[f_i] @ Dis_DB[curr_step + 1 - f_i - 1]
    for f_i in range(2, min(max_step, curr_step - 2) + 1):          # n-1 step
        rhs += G_DB[f_i] @ Trac_DB[curr_step - f_i - 1]
        rhs -= H_DB[f_i] @ Dis_DB[curr_step - f_i - 1]
    return G_curr, H_curr, rhs


def _matlab_march(p, numstep):
    K = O.max_step(p.E, p.nu, p.rho, p.coords, p.dt)
    num_matrix = min(K, numstep)
    G_DB, H_DB = _matlab_database(_per_step_blocks(p, num_matrix), num_matrix)
    N = p.ndof
    U = np.zeros((numstep, N))
    T = np.zeros((numstep, N))
    ubc = np.asarray(p.ubc_dof)
    for curr_step in range(2, numstep + 1):
        s = curr_step - 1
        y = np.zeros(N)                              # SetLoad with f = 1 (SolverT.cpp:344-377)
        y[p.fbc_dof] = p.fbc_val
        y[ubc] = p.ubc_val
        G, H, hist = _bem_averaging(G_DB, H_DB, U, T, curr_step, num_matrix + 1)
        G, H = G.copy(), H.copy()
        G[:, ubc], H[:, ubc] = -H[:, ubc], -G[:, ubc]   # ExchangeColumns (SolverT.cpp:265-342)
        x = np.linalg.solve(H, G @ y + hist)
        U[s] = x
        U[s, ubc] = p.ubc_val
        T[s] = 0.0
        T[s, p.fbc_dof] = p.fbc_val
        keep = ubc[ubc < N - 1]                      # fDofHigh inclusive (SolverStepLinearT.cpp:527)
        T[s, keep] = x[keep]
    return U, T


def test_literal_march_matches_the_matlab_restatement():
    p = helpers.small_bar(numstep=30)
    U, T = _matlab_march(p, p.numstep)
    K = O.max_step(p.E, p.nu, p.rho, p.coords, p.dt)
    MG, MH = O.build_lags(4, 1, 6, p.E, p.nu, p.rho, p.coords, p.ien, 0, p.dt, p.numstep, K)
    Ur, Tr = O.march(MG, MH, p.numstep, p.dt, p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val,
                     a1=0.25, a2=0.25)
    assert helpers.rel_fro(Ur, U) <= 1e-11
    assert helpers.rel_fro(Tr, T) <= 1e-11
    # the database of FE_BE_driver.m is UpdateGHLinear's lag store
    G_DB, H_DB = _matlab_database(_per_step_blocks(p, min(K, p.numstep)), min(K, p.numstep))
    for k in range(min(K, p.numstep) + 1):
        assert np.array_equal(G_DB[k + 1], MG[k]) or helpers.rel_fro(G_DB[k + 1], MG[k]) <= 1e-15
        assert np.array_equal(H_DB[k + 1], MH[k]) or helpers.rel_fro(H_DB[k + 1], MH[k]) <= 1e-15


def test_matlab_comparison_detects_a_changed_coefficient():
    """The comparison is not vacuous: the oracle run with a2 = 0.26 instead of 1/4 (which also
    leaves the (a2 - a1) M_0 x_{s-1} term that the Matlab form drops at a1 = a2) differs from
    the Matlab restatement by far more than the tolerance above."""
    p = helpers.small_bar(numstep=30)
    U, _ = _matlab_march(p, p.numstep)
    K = O.max_step(p.E, p.nu, p.rho, p.coords, p.dt)
    MG, MH = O.build_lags(4, 1, 6, p.E, p.nu, p.rho, p.coords, p.ien, 0, p.dt, p.numstep, K)
    Ur, _ = O.march(MG, MH, p.numstep, p.dt, p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val,
                    a1=0.25, a2=0.26)
    assert helpers.rel_fro(Ur, U) > 1e-4
