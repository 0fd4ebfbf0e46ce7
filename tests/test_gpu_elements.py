"""The element entry (tdbem_element_*; TDElementT::FormGH_Regular / FormGH_Singular /
GetGHLinear, src/TDElementT.h:37-52, and its Step, Heaviside and B-spline classes) against the
element matrices the reference's own classes wrote (tests/golden/elements.npz: 192 pairs, EType 1
and 2, regular and every singular node, steps 1, 2, 4, 7), against the oracle for the ramped
Heaviside kernel and the 10-point rule (no golden vectors), the device-array entry against the
host-array one, the host classes through tdbem_element, and the error paths."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O
from tdbem import api

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"
# Largest difference over a pair's 306 values, relative to its largest value, against a bound
# set by the pair's own conditioning.  Singular pairs at later steps cancel their Stokes and
# Kelvin parts, and moving a pair's coordinates by one ulp moves the oracle's own outputs by up to
# 4e-7 of their largest value (B-spline pair 191 of the golden set at step 7; 1e-13 to 1e-11 for
# most pairs).  The device evaluates the same integrals in another order, so its rounding
# differs by more than one such move: the bound is 50x the largest move over a few
# perturbations, at least FLOOR.  A wrong term, sign or node is O(1e-3) or more.
FLOOR = 5e-11


def _bound(kernel, etype, nq, E, nu, rho, coll, X, n, dt, sing, trials=8):
    ennd = 4 if etype == 1 else 8
    base = _flat(O.element(kernel, etype, nq, E, nu, rho, coll, X, n * dt, dt, n, sing), ennd)
    rng = np.random.default_rng(n * 1000 + sing + 7)
    ulp = np.spacing(np.abs(X).max())        # one ulp at the element's coordinate scale
    move = 0.0
    for _ in range(trials):
        Xp = X + rng.choice([-1.0, 1.0], X.shape) * ulp
        cp = Xp[sing].copy() if sing >= 0 else coll + rng.choice([-1.0, 1.0], 3) * ulp
        o = _flat(O.element(kernel, etype, nq, E, nu, rho, cp, Xp, n * dt, dt, n, sing), ennd)
        move = max(move, _rel(o, base))
    return max(FLOOR, 50 * move), base


def _flat(out, ennd):
    """GE_1, GE_2, HE_1, HE_2 padded to 24 columns, then C_1, C_2: the golden's layout."""
    mats = [np.pad(np.asarray(o).reshape(3, 3 * ennd), ((0, 0), (0, 24 - 3 * ennd))).ravel()
            for o in out[:4]]
    return np.concatenate(mats + [np.asarray(out[4]).ravel(), np.asarray(out[5]).ravel()])


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_element_matches_reference_golden_pairs(gpu):
    g = np.load(GOLD / "elements.npz")
    E, nu, rho, dt, nq = float(g["E"]), float(g["nu"]), float(g["rho"]), float(g["dt"]), int(g["nq"])
    print(f"[gpu case] element entry vs elements.npz ({len(g['out'])} pairs)", flush=True)
    groups = {}
    for q in range(len(g["out"])):
        groups.setdefault((int(g["kernel"][q]), int(g["etype"][q]), int(g["n"][q])), []).append(q)
    worst = 0.0
    for (kernel, etype, n), qs in sorted(groups.items()):
        ennd = 4 if etype == 1 else 8
        with api.Element(etype, kernel, nq, E=E, nu=nu, rho=rho) as el:
            # all pairs of one (class, type, step) in one call
            out = el.form(g["X"][qs][:, :ennd], g["coll"][qs], g["sing"][qs], n * dt, dt)
            for j, q in enumerate(qs):
                err = _rel(_flat([o[j] for o in out], ennd), g["out"][q])
                bound, _ = _bound(kernel, etype, nq, E, nu, rho, g["coll"][q], g["X"][q][:ennd],
                                  n, dt, int(g["sing"][q]))
                worst = max(worst, err / bound)
                assert err <= bound, (q, kernel, etype, n, int(g["sing"][q]), err, bound)
            # and one pair at a time, as the reference forms them
            q = qs[0]
            one = el.form(g["X"][q][:ennd], g["coll"][q], int(g["sing"][q]), n * dt, dt)
            assert np.array_equal(_flat(one, ennd), _flat([o[0] for o in out], ennd))
    print(f"largest error / bound {worst:.2f}")


def _random_pairs(rng, etype, count):
    """Perturbed unit quads (8-node: with mid-side nodes) and collocation points: regular ones at
    a distance, singular ones at a node."""
    ennd = 4 if etype == 1 else 8
    base = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], float)
    if etype == 2:
        base = np.vstack([base, [[.5, 0, 0], [1, .5, 0], [.5, 1, 0], [0, .5, 0]]])
    X = base[None] * rng.uniform(0.3, 1.5, (count, 1, 1)) + rng.normal(0, 0.04, (count, ennd, 3))
    sing = rng.integers(-1, ennd, count).astype(np.int32)
    coll = np.empty((count, 3))
    for p in range(count):
        if sing[p] >= 0:
            coll[p] = X[p, sing[p]]
        else:
            coll[p] = X[p].mean(0) + rng.normal(0, 1, 3) * rng.uniform(0.3, 3)
    return X, coll, sing


@pytest.mark.parametrize("kernel,etype,nq", [(6, 1, 6), (6, 2, 6), (4, 1, 10), (2, 2, 10)])
def test_element_matches_oracle(gpu, kernel, etype, nq):
    """The ramped-Heaviside element (no golden pairs) and the 10-point rule: the oracle's
    element routine, pinned by the golden pairs and, for the ramp, by the reference's ramp
    functions."""
    rng = np.random.default_rng(100 * kernel + 10 * etype + nq)
    E, nu, rho, dt = 2.0, 0.3, 1.5, 0.2
    X, coll, sing = _random_pairs(rng, etype, 24)
    ennd = X.shape[1]
    print(f"[gpu case] element entry vs oracle kernel {kernel} etype {etype} nq {nq}", flush=True)
    with api.Element(etype, kernel, nq, E=E, nu=nu, rho=rho) as el:
        for n in (1, 3, 8):
            out = el.form(X, coll, sing, n * dt, dt)
            for p in range(len(X)):
                bound, ref = _bound(kernel, etype, nq, E, nu, rho, coll[p], X[p], n, dt,
                                    int(sing[p]))
                err = _rel(_flat([o[p] for o in out], ennd), ref)
                assert err <= bound, (n, p, int(sing[p]), err, bound)


def test_element_device_entry_matches_host_entry(gpu):
    import torch
    rng = np.random.default_rng(5)
    X, coll, sing = _random_pairs(rng, 2, 300)
    sing[7] = 8                      # out of range: NaN from the device entry
    dev = torch.device("cuda:0")
    print("[gpu case] element device entry vs host entry", flush=True)
    with api.Element(2, api.KERNEL_STEP, 6, E=1.0, nu=0.25, rho=1.0) as el:
        tX = torch.from_numpy(X).to(dev)
        tc = torch.from_numpy(coll).to(dev)
        ts = torch.from_numpy(sing).to(dev)
        out = torch.empty(300 * (36 * 8 + 18), dtype=torch.float64, device=dev)
        el.form_device(tX, tc, ts, 0.3, 0.3, out)
        el.synchronize()
        got = out.cpu().numpy()
        ok = np.ones(300, bool)
        ok[7] = False
        host = el.form(X[ok], coll[ok], sing[ok], 0.3, 0.3)
        w = 300 * 72
        for b in range(6):
            blk = got[b * w:(b + 1) * w].reshape(300, 3, 24) if b < 4 else \
                got[4 * w + (b - 4) * 2700:4 * w + (b - 3) * 2700].reshape(300, 3, 3)
            assert np.array_equal(blk[ok], host[b]), b
            assert np.isnan(blk[7]).all()
        # the C block of the Step element is there at t == t_1 only
        assert np.abs(host[4]).max() > 0
        later = el.form(X[ok], coll[ok], sing[ok], 0.6, 0.3)
        assert not np.abs(later[4]).any() and not np.abs(later[5]).any()


def test_element_errors(gpu):
    print("[gpu case] element entry errors", flush=True)
    for bad in (dict(etype=3), dict(nq=7), dict(kernel=9), dict(nu=0.5)):
        args = dict(etype=1, kernel=api.KERNEL_STEP, nq=6, E=1.0, nu=0.25, rho=1.0)
        args.update(bad)
        with pytest.raises(api.TdbemError) as e:
            api.Element(args.pop("etype"), args.pop("kernel"), args.pop("nq"), **args)
        assert e.value.code == api.ERR_ARG
    with api.Element(1, api.KERNEL_STEP, 6, E=1.0, nu=0.25, rho=1.0) as el:
        X = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], float)
        for sing, t_1 in ((4, 0.1), (-2, 0.1), (0, 0.0)):
            with pytest.raises(api.TdbemError) as e:
                el.form(X, X[0], sing, 0.1, t_1)
            assert e.value.code == api.ERR_ARG
        # the element still works after a refused call
        out = el.form(X, X[0] + [0.5, 0.5, 1.0], -1, 0.1, 0.1)
        assert np.isfinite(out[0]).all()


def test_host_element_classes_match_golden(tmp_path):
    """TDElementStepLinearT / TDElementHeavisideLinearT / TDElementRampHeaviLinearT of the host
    library (tdbem_element: the G_H_Driver call sequence per pair) on the golden pairs."""
    g = np.load(GOLD / "elements.npz")
    E, nu, rho, dt, nq = float(g["E"]), float(g["nu"]), float(g["rho"]), float(g["dt"]), int(g["nq"])
    rec = np.dtype([("i", "<i4", 4), ("d", "<f8", 32)])
    recs = np.zeros(len(g["out"]), rec)
    for q in range(len(recs)):
        recs[q]["i"] = [g["kernel"][q], g["etype"][q], nq, g["sing"][q]]
        recs[q]["d"] = np.concatenate([[E, nu, rho, g["n"][q] * dt, dt], g["coll"][q],
                                       g["X"][q].ravel()])
    recs.tofile(tmp_path / "pairs.bin")
    exe = Path(api.PKG) / "tdbem_element"
    print("[gpu case] tdbem_element pairs.bin", flush=True)
    r = subprocess.run([str(exe), "pairs.bin", "out.bin"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(tmp_path / "out.bin").reshape(len(recs), 306)
    for q in range(len(recs)):
        etype = int(g["etype"][q])
        bound, _ = _bound(int(g["kernel"][q]), etype, nq, E, nu, rho, g["coll"][q],
                          g["X"][q][:4 if etype == 1 else 8], int(g["n"][q]), dt,
                          int(g["sing"][q]))
        assert _rel(got[q], g["out"][q]) <= bound, q
