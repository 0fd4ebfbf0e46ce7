"""Writes the golden fixtures of this directory from the reference's own compiled classes
(oracle/_ref/libtdbem_ref.so, built from /root/reference/src by `make -C oracle ref`).

Run in the development container, where the reference tree is mounted:
    python tests/golden/make_golden.py [green_ramp] [elements] [layer_decks]
Each fixture holds the inputs and the reference's outputs (layer_decks: the data of two input
decks, for the two-region solver, which needs PETSc in the reference); no reference source is
copied.
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parents[1] / "td-bem_amd"))

import helpers  # noqa: E402
import oracle_lib as O  # noqa: E402


def element_cases(seed=11):
    rng = np.random.default_rng(seed)
    X4 = np.array([[0, 0, 0], [1, 0, 0], [1.1, 0.9, 0.15], [0, 1, 0.05]], float)
    X8 = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0.1], [0, 1, 0], [0.5, 0.02, 0],
                   [1.02, 0.5, 0.05], [0.5, 1, 0.06], [0, 0.5, 0.01]], float)
    E, nu, rho = 1.0, 0.25, 1.0
    dt = 0.3
    rows = []
    for kern in (4, 2, 5):
        for etype, X in ((1, X4), (2, X8)):
            ennd = 4 if etype == 1 else 8
            colls = [(rng.uniform(-1, 2, 3), -1), (rng.uniform(-0.5, 1.5, 3) + [0, 0, 0.7], -1)]
            colls += [(X[a], a) for a in range(ennd)]
            for coll, sing in colls:
                for n in (1, 2, 4, 7):
                    out = O.element(kern, etype, 6, E, nu, rho, coll, X, n * dt, dt, n, sing,
                                    use_ref=True)
                    rows.append(dict(kernel=kern, etype=etype, coll=np.asarray(coll, float),
                                     X=np.pad(X, ((0, 8 - ennd), (0, 0))), n=n, sing=sing,
                                     out=np.concatenate([np.pad(o.reshape(3, -1),
                                                                ((0, 0), (0, 24 - o.reshape(3, -1).shape[1]))).ravel()
                                                         for o in out[:4]] + [out[4].ravel(), out[5].ravel()])))
    return rows


def green_ramp_cases(seed=17, count=120):
    """The ramped-Heaviside functions of GreenFunctionT at random points and times: the 40-point
    time integrals (DU1|DU2|DT1|DT2|K1|K2), Stokes_RampedHeaviside at a time between steps
    (U|T), and the Int_RampedHeaviside_Linear weights (g1|g2)."""
    rng = np.random.default_rng(seed)
    rows = []
    for _ in range(count):
        E, nu, rho = rng.uniform(0.5, 3), rng.uniform(0.0, 0.45), rng.uniform(0.5, 2)
        d = rng.normal(size=3) * rng.uniform(0.05, 3)
        nrm = rng.normal(size=3)
        dt = rng.uniform(0.01, 0.5)
        t = dt * rng.integers(1, 40)
        tp = t + dt * rng.uniform(-1, 1)
        lin = O.green(6, E, nu, rho, d, nrm, t, dt, use_ref=True)
        U, T, g = O.ramp_point(E, nu, rho, d, nrm, tp, dt, use_ref=True)
        rows.append((np.array([E, nu, rho, t, tp, dt]), d, nrm,
                     np.concatenate([x.ravel() for x in lin] + [U.ravel(), T.ravel(), g])))
    return rows


def write_layer_decks():
    """geometry/Layer_U_test0.txt and Layer_B_test0.txt, the two regions of the reference's
    two-region (SolverMRT) decks, as parsed by InputT::ReadInput: the deck data only."""
    from tdbem import deck
    ref_root = Path("/root/reference/geometry")
    out = {}
    for r, name in enumerate(("Layer_U_test0", "Layer_B_test0")):
        d = deck.read_deck(ref_root / f"{name}.txt")
        out.update({f"coords{r}": d.coords, f"ien{r}": d.ien, f"iface{r}": d.interface,
                    f"fbc_dof{r}": d.fbc_dof, f"fbc_val{r}": d.fbc_val,
                    f"scalars{r}": np.array([d.E, d.nu, d.density, d.dt, d.numstep, d.nq,
                                             d.etype, d.if_ext, d.interpolation], float)})
    np.savez_compressed(HERE / "layer_decks.npz", **out)


def main(names):
    if "layer_decks" in names:
        write_layer_decks()
    if O.ref() is None:
        raise SystemExit("oracle/_ref/libtdbem_ref.so is not built (make -C oracle ref)")
    if "green_ramp" in names:
        rows = green_ramp_cases()
        np.savez_compressed(HERE / "green_ramp.npz", scalars=np.array([r[0] for r in rows]),
                            direction=np.array([r[1] for r in rows]),
                            normal=np.array([r[2] for r in rows]),
                            out=np.array([r[3] for r in rows]))
    if "elements" in names or "assembly" in names:
        write_elements_and_assembly()
    print("golden fixtures written to", HERE)


def write_elements_and_assembly():
    rows = element_cases()
    np.savez_compressed(
        HERE / "elements.npz",
        kernel=np.array([r["kernel"] for r in rows], np.int32),
        etype=np.array([r["etype"] for r in rows], np.int32),
        coll=np.array([r["coll"] for r in rows]), X=np.array([r["X"] for r in rows]),
        n=np.array([r["n"] for r in rows], np.int32),
        sing=np.array([r["sing"] for r in rows], np.int32),
        out=np.array([r["out"] for r in rows]), E=1.0, nu=0.25, rho=1.0, dt=0.3, nq=6)

    # geometry/HalfSpacePlane_1E.txt and a coarse bar: G_H_Driver rows for the first steps
    for name, p, rows_hi in (("plane1e", helpers.plane_1e(), None),
                             ("bar_h05", helpers.small_bar(), 7)):
        upper = p.nnd - 1 if rows_hi is None else rows_hi
        steps = [1, 2, 3, 9]
        mats = []
        for n in steps:
            mats.append(np.stack(O.assemble_step(4, 1, 6, p.E, p.nu, p.rho, p.coords, p.ien,
                                                 p.if_ext, 0, upper, n * p.dt, p.dt, n,
                                                 use_ref=True)))
        np.savez_compressed(HERE / f"assemble_{name}.npz", coords=p.coords, ien=p.ien, E=p.E,
                            nu=p.nu, rho=p.rho, dt=p.dt, if_ext=p.if_ext, upper=upper,
                            steps=np.array(steps, np.int32), GH=np.array(mats))


if __name__ == "__main__":
    # python tests/golden/make_golden.py [green_ramp] [elements] [layer_decks]  (default: all)
    main(sys.argv[1:] or ["green_ramp", "elements", "layer_decks"])
