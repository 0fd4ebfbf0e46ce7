"""Pins the oracle's two-region solver (or_mrt_march, SolverMRT of src/SolverMRT.cpp).

PETSc is not installed, so the reference class cannot run here.  The restatement assembles the
global matrices through the reference's interior / interface index maps (InterfaceIndexing
:223-568, Assemble_GH_R1/R2 :685-850) and marches them with the loops of Solve_Direct
(:1035-1099) or Solve_Averaging (:1101-1249).  It is pinned by:
  - a region-wise formulation written here from the physics instead of the index maps: each
    region's own convolution sum_k (H_k u - G_k t) = 0, displacement continuity written as
    extra equations, traction equilibrium t_2 = -t_1 on the interface, and the averaging scheme
    written as one weighted (1, 2, 1) sum over neighbouring steps.  Any slip in a column map,
    the sign of the interface traction, a lag index or the load time fails it;
  - two regions without an interface reducing to two independent single-region marches of the
    (separately pinned) single-region oracle;
  - the interface-mismatch error and the closed-form peak of the reference's load pulse."""
import math

import numpy as np
import pytest

import oracle_lib as O


def _load(kind, t, amp=1.0, t_rise=0.0):
    if kind == 0:
        return 5.46366e9 * t * math.exp(-2000.0 * t)
    if t_rise > 0:
        return amp * min(1.0, t / t_rise)
    return amp


def _random_case(seed, nnd=(7, 6), iface1=(1, 4, 5), iface2=(0, 3, 2), perm=(2, 0, 1), K=4):
    """Two regions with random lag matrices.  Region 2's interface node iface2[i] sits where
    region 1's iface1[perm[i]] does, so the matching has to follow coordinates, not order."""
    rng = np.random.default_rng(seed)
    c1 = rng.standard_normal((nnd[0], 3))
    c2 = rng.standard_normal((nnd[1], 3)) + 10.0
    for i, j in enumerate(perm):
        c2[iface2[i]] = c1[iface1[j]]
    lags = []
    for r in range(2):
        N = 3 * nnd[r]
        MG = np.stack([rng.standard_normal((N, N)) * (0.4 / (k + 1)) for k in range(K + 1)])
        MH = np.stack([rng.standard_normal((N, N)) * (0.4 / (k + 1)) for k in range(K + 1)])
        MH[0] += 4.0 * np.eye(N)
        MG[0] += 1.5 * np.eye(N)
        lags.append((MG, MH))
    # loads: interior dofs plus one interface dof, which SolverMRT does not load
    regions = [
        dict(coords=c1, iface=np.array(iface1, dtype=int),
             fbc_dof=np.array([0, 2, 3 * iface1[0] + 1 if iface1 else 5, 7]),
             fbc_val=np.array([1.0, -0.5, 9.0, 0.25])),
        dict(coords=c2, iface=np.array(iface2, dtype=int),
             fbc_dof=np.array([3 * 1 + 2, 3 * 4, 3 * iface2[1] if iface2 else 9]),
             fbc_val=np.array([0.75, -1.25, 4.0])),
    ]
    return regions, lags


def _regionwise(regions, lags, numstep, dt, K, scheme, loadf):
    """Region-wise restatement: unknowns z = [u_1, u_2, tau] with tau the interface traction of
    region 1 (region 2 carries -tau), equations = the two regions' step equations plus
    continuity u_1 = u_2 on the interface.  Returns U_1, U_2 [numstep+1][3 nnd] and tau."""
    c1, c2 = regions[0]["coords"], regions[1]["coords"]
    i1, i2 = regions[0]["iface"], regions[1]["iface"]
    nz = len(i1)
    # region 2 interface node i sits at region 1 interface slot slot[i]
    slot = [int(np.argmin(np.linalg.norm(c1[i1] - c2[n], axis=1))) for n in i2]
    N = [3 * len(c1), 3 * len(c2)]
    iface_dofs = [np.array([3 * n + d for n in ifc for d in range(3)], dtype=int) for ifc in (i1, i2)]
    # traction of region r = lv_r f + P_r tau (P_1 = +selection, P_2 = -selection in slot order)
    lv, P = [], []
    for r in range(2):
        v = np.zeros(N[r])
        isif = np.zeros(N[r], bool)
        isif[iface_dofs[r]] = True
        for d, val in zip(regions[r]["fbc_dof"], regions[r]["fbc_val"]):
            if not isif[d]:
                v[d] = val
        lv.append(v)
        p = np.zeros((N[r], 3 * nz))
        for q, n in enumerate(regions[r]["iface"]):
            sl = q if r == 0 else slot[q]
            for d in range(3):
                p[3 * n + d, 3 * sl + d] = 1.0 if r == 0 else -1.0
        P.append(p)
    nu = N[0] + N[1]
    U = [np.zeros((numstep + 1, N[0])), np.zeros((numstep + 1, N[1]))]
    Tr = [np.zeros((numstep + 1, N[0])), np.zeros((numstep + 1, N[1]))]
    tau = np.zeros((numstep + 1, 3 * nz))

    def lagsum(r, k, weights_steps):
        MG, MH = lags[r]
        acc = np.zeros(N[r])
        for w, m in weights_steps:
            if m >= 1 and k <= K:
                acc += w * (MG[k] @ Tr[r][m] - MH[k] @ U[r][m])
        return acc

    def solve(s, c0, c1, f, hist):
        A = np.zeros((nu + 3 * nz, nu + 3 * nz))
        b = np.zeros(nu + 3 * nz)
        row = 0
        for r in range(2):
            MG, MH = lags[r]
            Gc = c0 * MG[0] + c1 * MG[1]
            Hc = c0 * MH[0] + c1 * MH[1]
            cols = slice(0, N[0]) if r == 0 else slice(N[0], nu)
            A[row:row + N[r], cols] = Hc
            A[row:row + N[r], nu:] = -Gc @ P[r]
            b[row:row + N[r]] = Gc @ lv[r] * f + hist[r]
            row += N[r]
        for q in range(nz):
            for d in range(3):
                A[row, 3 * i1[q] + d] = 1.0
                A[row, N[0] + 3 * i2[slot.index(q)] + d] = -1.0
                row += 1
        z = np.linalg.solve(A, b)
        U[0][s], U[1][s], tau[s] = z[:N[0]], z[N[0]:nu], z[nu:]
        for r in range(2):
            Tr[r][s] = lv[r] * f + P[r] @ tau[s]

    if scheme == 0:
        for n in range(1, numstep + 1):
            hist = [sum((lagsum(r, k, [(1.0, n - k)]) for k in range(1, min(n, K) + 1)),
                        np.zeros(N[r])) for r in range(2)]
            solve(n, 1.0, 0.0, loadf(n * dt), hist)
    else:
        for s in range(1, numstep):
            f = loadf((s + 1) * dt)
            if s == 1:
                solve(1, 1.0, 0.0, f, [np.zeros(N[0]), np.zeros(N[1])])
                continue
            # sum_k A_k (x_{s+1-k} + 2 x_{s-k} + x_{s-1-k}) with every k = 0 step taken as x_s
            hist = []
            for r in range(2):
                h = lagsum(r, 1, [(2.0, s - 1), (1.0, s - 2)])
                for k in range(2, K + 1):
                    h += lagsum(r, k, [(1.0, s + 1 - k), (2.0, s - k), (1.0, s - 1 - k)])
                hist.append(h)
            solve(s, 4.0, 1.0, f, hist)
    return U, tau


@pytest.mark.parametrize("scheme", [0, 1])
@pytest.mark.parametrize("load_kind", [0, 1])
def test_mrt_matches_regionwise_formulation(scheme, load_kind):
    K, numstep, dt = 4, 9, 2e-4
    regions, lags = _random_case(11 + scheme)
    X, Ureg = O.mrt_march(regions, lags, numstep, dt, K, scheme=scheme, load_kind=load_kind,
                          amp=2.0, t_rise=3.5 * dt)
    loadf = lambda t: _load(load_kind, t, 2.0, 3.5 * dt)
    Uw, tau = _regionwise(regions, lags, numstep, dt, K, scheme, loadf)
    scale = max(np.abs(Uw[0]).max(), np.abs(Uw[1]).max(), 1e-300)
    for r in range(2):
        np.testing.assert_allclose(Ureg[r], Uw[r], rtol=0, atol=1e-11 * scale)
    # the last 3 n_iface unknowns are the interface traction of region 1
    nz3 = 3 * len(regions[0]["iface"])
    np.testing.assert_allclose(X[:, -nz3:], tau, rtol=0, atol=1e-11 * np.abs(tau).max())
    assert np.abs(Uw[0][2:]).max() > 0


@pytest.mark.parametrize("scheme", [0, 1])
def test_mrt_without_interface_is_two_single_region_marches(scheme):
    """No interface: each region is a pure-traction single-region problem, which the single-
    region oracle marches with alpha1 = alpha2 = 0 (Solve_Direct) or 1/4 (Solve_Averaging; the
    load is held constant because SolverMRT's averaging loads step s at (s + 1) dt)."""
    K, numstep, dt = 3, 8, 1e-3
    regions, lags = _random_case(5, iface1=(), iface2=(), perm=(), K=K)
    t_rise = 4.5 * dt if scheme == 0 else 0.0
    X, Ureg = O.mrt_march(regions, lags, numstep, dt, K, scheme=scheme, load_kind=1, amp=1.5,
                          t_rise=t_rise)
    a = 0.0 if scheme == 0 else 0.25
    for r in range(2):
        MG, MH = lags[r]
        U, T = O.march(MG, MH, numstep + 1 if scheme == 0 else numstep, dt, [], [],
                       regions[r]["fbc_dof"], regions[r]["fbc_val"], a1=a, a2=a, amp=1.5,
                       t_rise=t_rise)
        ref = U if scheme == 0 else np.vstack([U, np.zeros((1, U.shape[1]))])
        np.testing.assert_allclose(Ureg[r], ref, rtol=0, atol=1e-11 * np.abs(ref).max())


def test_mrt_interface_mismatch():
    regions, lags = _random_case(3)
    regions[1]["coords"] = regions[1]["coords"].copy()
    regions[1]["coords"][regions[1]["iface"][1], 2] += 1e-6
    with pytest.raises(ValueError, match="Interface mismatch"):
        O.mrt_march(regions, lags, 4, 1e-3, 4)


def test_mrt_load_pulse_peak():
    """f(t) = 5.46366e9 t exp(-2000 t) (src/SolverMRT.cpp:937-939) peaks at t = 1/2000 with
    f = 5.46366e9 / (2000 e) = 1.005e6, i.e. a 1 MPa pulse."""
    f = O.oracle().or_mrt_load_factor
    cfg = O.MrtCfg()
    cfg.load_kind = 0
    ts = np.linspace(0, 2e-3, 4001)
    vals = np.array([f(O.C.byref(cfg), O.D(t)) for t in ts])
    assert abs(ts[np.argmax(vals)] - 5e-4) <= 1e-6
    assert vals.max() == pytest.approx(5.46366e9 / (2000 * math.e), rel=1e-9)
    assert vals.max() == pytest.approx(1.005e6, rel=1e-3)


def test_split_bar_behaves_like_the_uncut_bar():
    """Physics: a free bar cut in two by an interface (the two-region solver on real lag
    matrices) moves like the uncut bar (the single-region oracle, same alpha = 1/4 averaging,
    constant load so the load-time shift of Solve_Averaging does not matter).  Mean u_z of the
    loaded top face and of the free bottom face agree within 10 % (4 % r.m.s.) once they are
    larger than 0.5; the wave reaches the bottom at about step 16 in both."""
    from tdbem import configs, meshgen
    r1, r2, i1, i2 = configs.split_bar(h=0.5, numstep=40)
    E, nu, rho, dt, numstep = r1.E, r1.nu, r1.rho, r1.dt, r1.numstep
    K = max(O.max_step(E, nu, rho, r.coords, dt) for r in (r1, r2))
    lags = [O.build_lags(4, 1, 6, E, nu, rho, r.coords, r.ien, 0, dt, numstep, K) for r in (r1, r2)]
    regions = [dict(coords=r1.coords, iface=i1, fbc_dof=r1.fbc_dof, fbc_val=r1.fbc_val),
               dict(coords=r2.coords, iface=i2, fbc_dof=r2.fbc_dof, fbc_val=r2.fbc_val)]
    X, U = O.mrt_march(regions, lags, numstep, dt, K, scheme=1, load_kind=1, amp=1.0)
    mb = meshgen.box_surface(1.0, 1.0, 4.0, 0.5)
    Kb = O.max_step(E, nu, rho, mb.coords, dt)
    MG, MH = O.build_lags(4, 1, 6, E, nu, rho, mb.coords, mb.ien, 0, dt, numstep, Kb)
    topb, botb = mb.tags["ztop"], mb.tags["z0"]
    Ub, _ = O.march(MG, MH, numstep, dt, [], [], 3 * topb + 2, np.ones(len(topb)), a1=0.25,
                    a2=0.25)
    top2 = r2.fbc_dof   # z dofs of region 2's top face
    bot1 = 3 * meshgen.box_surface(1.0, 1.0, 2.0, 0.5).tags["z0"] + 2
    pairs = [(U[1][:numstep, top2].mean(1), Ub[:, 3 * topb + 2].mean(1)),
             (U[0][:numstep, bot1].mean(1), Ub[:, 3 * botb + 2].mean(1))]
    for got, ref in pairs:
        m = np.abs(ref) > 0.5
        assert m.sum() >= 10
        rel = np.abs(got[m] - ref[m]) / np.abs(ref[m])
        print("split bar: max rel", rel.max(), "rms rel", np.sqrt(np.mean(rel ** 2)))
        assert rel.max() < 0.1 and np.sqrt(np.mean(rel ** 2)) < 0.04


@pytest.mark.parametrize("scheme", [0, 1])
def test_region_order_does_not_matter(scheme):
    """Swapping the regions permutes the unknowns and flips the sign of the interface traction
    (tau is always region 1's traction); the displacements stay the same."""
    K, numstep, dt = 4, 8, 3e-4
    regions, lags = _random_case(21 + scheme)
    X, U = O.mrt_march(regions, lags, numstep, dt, K, scheme=scheme, load_kind=1, amp=1.0,
                       t_rise=2.5 * dt)
    Xs, Us = O.mrt_march(regions[::-1], lags[::-1], numstep, dt, K, scheme=scheme, load_kind=1,
                         amp=1.0, t_rise=2.5 * dt)
    scale = max(np.abs(U[0]).max(), np.abs(U[1]).max())
    np.testing.assert_allclose(Us[0], U[1], rtol=0, atol=1e-11 * scale)
    np.testing.assert_allclose(Us[1], U[0], rtol=0, atol=1e-11 * scale)
    # tau in region 2's interface order, as region 1 of the swapped run sees it
    i1, i2 = regions[0]["iface"], regions[1]["iface"]
    c1, c2 = regions[0]["coords"], regions[1]["coords"]
    nz = len(i1)
    tau = X[:, -3 * nz:].reshape(-1, nz, 3)
    taus = Xs[:, -3 * nz:].reshape(-1, nz, 3)
    for q, n in enumerate(i2):
        sl = int(np.argmin(np.linalg.norm(c1[i1] - c2[n], axis=1)))
        np.testing.assert_allclose(taus[:, q], -tau[:, sl], rtol=0,
                                   atol=1e-11 * np.abs(tau).max())
