"""UpdateGHLinear with caller-supplied blocks (tdbem_update_gh_linear; the reference's
SolverStepLinearT::UpdateGHLinear, src/SolverStepLinearT.cpp:58-97, called from main.cpp:121):
the oracle's own per-step G/H blocks fed step by step, interleaved with the march as main.cpp
interleaves them, reproduce the oracle's march; feeding all steps first gives the same histories
bit for bit; the driver's --caller-lags run (G_H_Driver's blocks handed to the solver) matches
its own-assembly run; and the call-order and symmetry errors."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import helpers
import oracle_lib as O
from tdbem import api, deck

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _blocks(p, n):
    """The oracle's G_H_Driver blocks of step n: (phi_1, phi_2) parts of G and H."""
    return O.assemble_step(4, p.etype, p.nq, p.E, p.nu, p.rho, p.coords, p.ien, p.if_ext, 0,
                           p.nnd - 1, n * p.dt, p.dt, n)


def _reference(p):
    K = O.max_step(p.E, p.nu, p.rho, p.coords, p.dt)
    MG, MH = O.build_lags(4, p.etype, p.nq, p.E, p.nu, p.rho, p.coords, p.ien, p.if_ext, p.dt,
                          p.numstep, K)
    return O.march(MG, MH, p.numstep, p.dt, p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val)


def _interleaved(p):
    """main.cpp:89-128: per step n, the blocks (n <= MaxStep), then the solve of x_{n-1}."""
    bc = (p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val)
    with api.Context.from_problem(p) as ctx:
        K = ctx.max_step
        started = False
        for n in range(1, p.numstep + 1):
            if n <= K:
                G1, G2, H1, H2 = _blocks(p, n)
                ctx.update_gh_linear(n, G1, G2, H1, H2)
            if n >= 2:
                if not started:
                    ctx.march_init(p.a1, p.a2, *bc)
                    started = True
                ctx.march_run(1)
        assert ctx.step_index == p.numstep - 1
        return ctx.history()


@pytest.mark.parametrize("case", ["bar", "sphere"])
def test_caller_blocks_interleaved_reproduce_the_oracle_march(gpu, case):
    p = {"bar": lambda: helpers.small_bar(numstep=30),
         "sphere": lambda: helpers.small_sphere(numstep=24)}[case]()
    print(f"[gpu case] caller-supplied blocks, interleaved, {case}", flush=True)
    U, T = _interleaved(p)
    Ur, Tr = _reference(p)
    eu, et = helpers.rel_fro(U, Ur), helpers.rel_fro(T, Tr)
    print(f"{case}: u {eu:.2e} t {et:.2e}")
    assert eu <= TOL and et <= TOL


def test_caller_blocks_all_first(gpu):
    """All steps' blocks before the march: the store holds the oracle's lags exactly, and the
    histories equal the interleaved run's to the rounding of summation order (the interleaved
    run's stores hold fewer lags per pair while later blocks are missing)."""
    p = helpers.small_bar(numstep=20)
    print("[gpu case] caller-supplied blocks, all steps first", flush=True)
    U0, T0 = _interleaved(p)
    with api.Context.from_problem(p) as ctx:
        for n in range(1, ctx.max_step + 1):
            G1, G2, H1, H2 = _blocks(p, n)
            ctx.update_gh_linear(n, G1, G2, H1, H2)   # steps past numstep are ignored
        ctx.march_init(p.a1, p.a2, p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val)
        ctx.march_run()
        U, T = ctx.history()
        # the lags are the oracle's accumulation of those blocks, exactly
        K = O.max_step(p.E, p.nu, p.rho, p.coords, p.dt)
        MG, MH = O.build_lags(4, 1, 6, p.E, p.nu, p.rho, p.coords, p.ien, 0, p.dt, p.numstep, K)
        for k in range(min(K, p.numstep) + 1):
            G, H = ctx.get_lag(k)
            assert np.array_equal(G, MG[k]) and np.array_equal(H, MH[k]), k
    assert helpers.rel_fro(U, U0) <= 1e-13 and helpers.rel_fro(T, T0) <= 1e-13


def test_caller_blocks_order_and_symmetry_errors(gpu):
    p = helpers.small_bar(numstep=10, h=1.0)
    bc = (p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val)
    print("[gpu case] caller-supplied blocks: errors", flush=True)
    with api.Context.from_problem(p) as ctx:
        G1, G2, H1, H2 = _blocks(p, 1)
        ctx.update_gh_linear(1, G1, G2, H1, H2)
        # the step-0 matrices need lag 1 complete, i.e. step 2's blocks
        with pytest.raises(api.TdbemError) as e:
            ctx.march_init(p.a1, p.a2, *bc)
        assert e.value.code == api.ERR_STATE
        G1, G2, H1, H2 = _blocks(p, 2)
        ctx.update_gh_linear(2, G1, G2, H1, H2)
        ctx.march_init(p.a1, p.a2, *bc)
        ctx.march_run(1)                       # x_1 needs lags 0..1
        with pytest.raises(api.TdbemError) as e:
            ctx.march_run(1)                   # x_2 needs lag 2, complete after step 3
        assert e.value.code == api.ERR_STATE
        G1, G2, H1, H2 = _blocks(p, 3)
        Gbad = G1.copy()
        Gbad[0, 4] += 1.0                      # node block (0, 1): G[0][1] != G[1][0]
        with pytest.raises(api.TdbemError) as e:
            ctx.update_gh_linear(3, Gbad, G2, H1, H2)
            ctx.march_run(1)
        assert e.value.code == api.ERR_ARG
    with api.Context.from_problem(p) as ctx:
        ctx.set_precision(api.PRECISION_MIXED)
        G1, G2, H1, H2 = _blocks(p, 1)
        with pytest.raises(api.TdbemError) as e:
            ctx.update_gh_linear(1, G1, G2, H1, H2)
        assert e.value.code == api.ERR_ARG


def test_driver_caller_lags_matches_own_assembly(tmp_path):
    p = helpers.small_bar(numstep=14)
    deck.write_deck(tmp_path / "bar.txt", p.coords, p.ien, E=p.E, nu=p.nu, density=p.rho,
                    dt=p.dt, numstep=p.numstep, ubc_dof=p.ubc_dof, ubc_val=p.ubc_val,
                    fbc_dof=p.fbc_dof, fbc_val=p.fbc_val)
    (tmp_path / "alpha.txt").write_text("0.25 0.25\n")
    exe = Path(api.PKG) / "tdbem_main"
    hist = {}
    for tag, extra in (("own", []), ("caller", ["--caller-lags"])):
        print(f"tdbem_main bar.txt --history {extra}", flush=True)
        out = subprocess.run([str(exe), "bar.txt", "--history", f"{tag}.bin"] + extra,
                             cwd=tmp_path, capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr
        hist[tag] = np.fromfile(tmp_path / f"{tag}.bin").reshape(2, p.numstep, 3 * p.nnd)
    Ur, Tr = _reference(p)
    for tag in hist:
        assert helpers.rel_fro(hist[tag][0], Ur) <= TOL, tag
        assert helpers.rel_fro(hist[tag][1], Tr) <= TOL, tag
