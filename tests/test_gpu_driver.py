"""The command-line driver tdbem_main (the step loop of src/main.cpp:22-149 over the host classes
of td-bem_amd/host) on the GPU, against the oracle: displacement.txt (RecordResult,
src/SolverStepLinearT.cpp:549-603), the full-precision --history, the FE-BE files of --write-gh
(SolverT::WriteGH, src/SolverT.cpp:670-732, with main.cpp:113-116's G1/G2 naming: file G1 holds
the phi_2 part fG_2, file G2 the phi_1 part fG_1) and their binary variant."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import helpers
import oracle_lib as O
from tdbem import api, deck

pytestmark = pytest.mark.gpu

TOL = 1e-10


def test_driver_single_region(tmp_path):
    p = helpers.small_bar(numstep=12)
    deck.write_deck(tmp_path / "bar.txt", p.coords, p.ien, E=p.E, nu=p.nu, density=p.rho,
                    dt=p.dt, numstep=p.numstep, ubc_dof=p.ubc_dof, ubc_val=p.ubc_val,
                    fbc_dof=p.fbc_dof, fbc_val=p.fbc_val)
    (tmp_path / "alpha.txt").write_text("0.25 0.25\n")
    exe = Path(api.PKG) / "tdbem_main"
    print("tdbem_main bar.txt --write-gh --write-gh-binary --history", flush=True)
    out = subprocess.run([str(exe), "bar.txt", "--write-gh", "--write-gh-binary", "--history",
                          "hist.bin"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "The computation terminated properly" in out.stdout
    N, S = 3 * p.nnd, p.numstep

    K = O.max_step(p.E, p.nu, p.rho, p.coords, p.dt)
    MG, MH = O.build_lags(4, 1, 6, p.E, p.nu, p.rho, p.coords, p.ien, 0, p.dt, S, K)
    Ur, Tr = O.march(MG, MH, S, p.dt, p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val, 0.25, 0.25)
    hist = np.fromfile(tmp_path / "hist.bin", dtype=np.float64).reshape(2, S, N)
    assert helpers.rel_fro(hist[0], Ur) <= TOL
    assert helpers.rel_fro(hist[1], Tr) <= TOL

    # displacement.txt: header, then (dt (n-1), load, u[2], t[2]) of x_{n-1}, 6 digits
    lines = (tmp_path / "displacement.txt").read_text().splitlines()
    assert lines[0].split() == ["%time", "load", "displacement", "traction"]
    rows = np.array([[float(v) for v in l.split()] for l in lines[1:]])
    assert rows.shape == (S, 4)
    np.testing.assert_allclose(rows[:, 0], p.dt * np.arange(S), rtol=1e-5)
    np.testing.assert_allclose(rows[:, 1], 1.0)
    for col, ref in ((2, Ur[:, 2]), (3, Tr[:, 2])):
        np.testing.assert_allclose(rows[:, col], ref, rtol=1e-5, atol=1e-6 * np.abs(ref).max())

    # FE-BE files for every assembled step (curr_step <= MaxStep), named by curr_step - 1
    steps = min(K, S)
    for n in (1, 2, steps):
        ref = O.assemble_step(4, 1, 6, p.E, p.nu, p.rho, p.coords, p.ien, 0, 0, p.nnd - 1,
                              n * p.dt, p.dt, n)
        for name, m in (("G1", ref[1]), ("G2", ref[0]), ("H1", ref[3]), ("H2", ref[2])):
            raw = np.fromfile(tmp_path / f"{name}_step{n - 1}.bin", dtype=np.int64, count=2)
            assert tuple(raw) == (N, N)
            b = np.fromfile(tmp_path / f"{name}_step{n - 1}.bin", dtype=np.float64,
                            offset=16).reshape(N, N)
            assert helpers.rel_fro(b, m) <= TOL, (name, n)
            txt = np.loadtxt(tmp_path / f"{name}_step{n - 1}.txt")
            assert txt.shape == (N, N)
            # %15.5e: 6 significant digits of the same numbers
            np.testing.assert_allclose(txt, b, rtol=6e-6, atol=1e-300)
    assert not (tmp_path / f"G1_step{steps}.txt").exists()
