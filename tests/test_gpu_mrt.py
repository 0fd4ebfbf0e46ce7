"""GPU parity of the two-region solver (tdbem_mrt_*; SolverMRT of src/SolverMRT.cpp) with the
oracle's restatement (or_mrt_march) on the oracle's own lag matrices, through the C ABI.

Cases: the reference's two-region decks geometry/Layer_U_test0.txt + Layer_B_test0.txt (their
data as the fixture layer_decks.npz; 50 + 25 nodes, 25 interface nodes, numstep 10, the
reference's fMaxStep 10 and its load pulse) under both schemes, and the free bar cut in two
(configs.split_bar) with a ramp load.  Bar: relative Frobenius error <= 1e-10 of the global
unknowns, the region displacements and tractions.  Solve_Direct on the layer decks grows by
six orders of magnitude in ten steps (collocation with beta = cs dt / h this large is
unstable; the averaging scheme is not; so is the bar's), so for Solve_Direct the bar is 1e-10
of each step's own norm."""
from pathlib import Path

import numpy as np
import pytest

import helpers
import oracle_lib as O
from tdbem import api, configs

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden"
TOL = 1e-10


def layer_regions():
    z = np.load(GOLD / "layer_decks.npz")
    regs = []
    for r in range(2):
        E, nu, rho, dt, numstep, nq, etype, if_ext, interp = z[f"scalars{r}"]
        assert int(interp) == 4   # the Step kernel
        regs.append(dict(coords=z[f"coords{r}"], ien=z[f"ien{r}"], iface=z[f"iface{r}"],
                         fbc_dof=z[f"fbc_dof{r}"], fbc_val=z[f"fbc_val{r}"], E=float(E),
                         nu=float(nu), rho=float(rho), dt=float(dt), numstep=int(numstep),
                         nq=int(nq), etype=int(etype), if_ext=int(if_ext)))
    return regs


def split_bar_regions(numstep):
    r1, r2, i1, i2 = configs.split_bar(h=0.5, numstep=numstep)
    return [dict(coords=p.coords, ien=p.ien, iface=i, fbc_dof=p.fbc_dof, fbc_val=p.fbc_val,
                 E=p.E, nu=p.nu, rho=p.rho, dt=p.dt, numstep=p.numstep, nq=p.nq, etype=p.etype,
                 if_ext=p.if_ext) for p, i in ((r1, i1), (r2, i2))]


def contexts(regs):
    return [api.Context(r["coords"], r["ien"], E=r["E"], nu=r["nu"], rho=r["rho"], dt=r["dt"],
                        numstep=r["numstep"], nq=r["nq"], etype=r["etype"], if_ext=r["if_ext"])
            for r in regs]


def run_gpu(regs, K, scheme, load_kind, amp=1.0, t_rise=0.0, chunks=None):
    ctx = contexts(regs)
    s = api.TwoRegionSolver(ctx[0], ctx[1], regs[0]["iface"], regs[1]["iface"], max_step=K)
    s.init(scheme, load_kind, amp, t_rise, fbc1=(regs[0]["fbc_dof"], regs[0]["fbc_val"]),
           fbc2=(regs[1]["fbc_dof"], regs[1]["fbc_val"]))
    for n in chunks or [None]:
        s.run(n)
    X, U, T = s.history()
    s.close()
    return X, U, T


def run_oracle(regs, K, scheme, load_kind, amp=1.0, t_rise=0.0):
    lags = [O.build_lags(4, r["etype"], r["nq"], r["E"], r["nu"], r["rho"], r["coords"],
                         r["ien"], r["if_ext"], r["dt"], r["numstep"], K) for r in regs]
    X, U = O.mrt_march(regs, lags, regs[0]["numstep"], regs[0]["dt"], K, scheme=scheme,
                       load_kind=load_kind, amp=amp, t_rise=t_rise)
    return X, np.hstack(U)


def expected_tractions(regs, X, f_of_step):
    """Region tractions from the global unknowns: the interior loads times f, +-tau on the
    interface (region 1's interface order, region 2's nodes matched by coordinates)."""
    nz = len(regs[0]["iface"])
    n0 = 3 * len(regs[0]["coords"])
    tau = X[:, -3 * nz:]
    T = np.zeros_like(X)
    for r, reg in enumerate(regs):
        base = 0 if r == 0 else n0
        isif = np.zeros(3 * len(reg["coords"]), bool)
        for q, n in enumerate(reg["iface"]):
            isif[3 * n:3 * n + 3] = True
            if r == 0:
                sl = q
            else:
                sl = int(np.argmin(np.linalg.norm(regs[0]["coords"][regs[0]["iface"]]
                                                  - reg["coords"][n], axis=1)))
            T[:, base + 3 * n:base + 3 * n + 3] = (1 if r == 0 else -1) * tau[:, 3 * sl:3 * sl + 3]
        for d, v in zip(reg["fbc_dof"], reg["fbc_val"]):
            if not isif[d]:
                T[:, base + d] = v * np.array([f_of_step(s) for s in range(X.shape[0])])
    return T


def check(got, ref, per_step=False):
    if per_step:
        for s in range(ref.shape[0]):
            nr = np.linalg.norm(ref[s])
            assert np.linalg.norm(got[s] - ref[s]) <= TOL * max(nr, 1e-300), f"step {s}"
    else:
        assert helpers.rel_fro(got, ref) <= TOL


@pytest.mark.parametrize("order", ["UB", "BU"])
@pytest.mark.parametrize("scheme", [api.MRT_DIRECT, api.MRT_AVERAGING])
def test_layer_decks(scheme, order):
    regs = layer_regions()
    if order == "BU":   # Layer_B as region 1: the order SolverMRT::RecordResult's dof implies
        regs = regs[::-1]
    print(f"layer decks {order}, scheme {scheme}", flush=True)
    X, U, T = run_gpu(regs, 10, scheme, api.MRT_LOAD_PULSE)
    Xr, Ur = run_oracle(regs, 10, scheme, 0)
    assert np.abs(Xr).max() > 0
    check(X, Xr, per_step=True)
    check(U, Ur, per_step=True)
    dt = regs[0]["dt"]
    f = lambda s: 0.0 if s == 0 else 5.46366e9 * (s + scheme) * dt * np.exp(-2000 * (s + scheme) * dt)
    Te = expected_tractions(regs, Xr, f)
    if scheme == 1:
        Te[-1] = 0.0   # x_numstep is never solved by Solve_Averaging
    check(T, Te, per_step=True)


@pytest.mark.parametrize("scheme,numstep", [(api.MRT_AVERAGING, 40), (api.MRT_DIRECT, 18)])
def test_split_bar(scheme, numstep):
    regs = split_bar_regions(numstep)
    ctx = contexts(regs)
    K = max(c.max_step for c in ctx)
    del ctx
    print(f"split bar, scheme {scheme}, K {K}", flush=True)
    t_rise = 6.5 * regs[0]["dt"]
    X, U, T = run_gpu(regs, K, scheme, api.MRT_LOAD_RAMP, amp=1.5, t_rise=t_rise,
                      chunks=[3, 1, 100])
    Xr, Ur = run_oracle(regs, K, scheme, 1, amp=1.5, t_rise=t_rise)
    check(X, Xr, per_step=scheme == api.MRT_DIRECT)
    check(U, Ur, per_step=scheme == api.MRT_DIRECT)


def test_interface_mismatch_is_an_argument_error():
    regs = layer_regions()
    regs[1]["coords"] = regs[1]["coords"].copy()
    regs[1]["coords"][regs[1]["iface"][3], 0] += 1e-6
    ctx = contexts(regs)
    with pytest.raises(api.TdbemError) as e:
        api.TwoRegionSolver(ctx[0], ctx[1], regs[0]["iface"], regs[1]["iface"])
    assert e.value.code == api.ERR_ARG
    assert "Interface mismatch" in e.value.msg


def test_driver_two_region_run(tmp_path):
    """tdbem_main B.txt --mrt U.txt: the step loop of the host SolverMRT class; displacement.txt
    holds (dt (n-1), region 2's dof 46*3-1 at step n-1) per step as SolverMRT::RecordResult
    writes it, and BEM_2nd_<n>.vtk the region-2 field of step n-1."""
    import subprocess
    from tdbem import deck
    regs = layer_regions()[::-1]
    paths = []
    for r, name in zip(regs, ("B.txt", "U.txt")):
        deck.write_deck(tmp_path / name, r["coords"], r["ien"], E=r["E"], nu=r["nu"],
                        density=r["rho"], dt=r["dt"], numstep=r["numstep"], nq=r["nq"],
                        if_ext=r["if_ext"], fbc_dof=r["fbc_dof"], fbc_val=r["fbc_val"],
                        formulation=3, interface=r["iface"])
        paths.append(str(tmp_path / name))
    exe = Path(api.PKG) / "tdbem_main"
    out = subprocess.run([str(exe), paths[0], "--mrt", paths[1]], cwd=tmp_path,
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rows = np.loadtxt(tmp_path / "displacement.txt")
    Xr, Ur = run_oracle(regs, 10, 0, 0)
    n0 = 3 * len(regs[0]["coords"])
    numstep = regs[0]["numstep"]
    assert rows.shape == (numstep, 2)
    np.testing.assert_allclose(rows[:, 0], regs[0]["dt"] * np.arange(numstep), rtol=1e-5)
    ref = Ur[:numstep, n0 + 46 * 3 - 1]
    np.testing.assert_allclose(rows[:, 1], ref, rtol=2e-5, atol=1e-6 * np.abs(ref).max())
    vtk = (tmp_path / f"BEM_2nd_{numstep}.vtk").read_text().split("\n")
    k = vtk.index("VECTORS Displacement double")
    field = np.array([[float(v) for v in line.split()] for line in vtk[k + 1:k + 1 + len(regs[1]["coords"])]])
    np.testing.assert_allclose(field.ravel(), Ur[numstep - 1, n0:], rtol=2e-5,
                               atol=1e-6 * np.abs(Ur[numstep - 1, n0:]).max())
