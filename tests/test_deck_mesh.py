"""Input-deck format (src/InputT.cpp:84-994) and the synthetic workload meshes."""
from pathlib import Path

import numpy as np
import pytest

from tdbem import configs, deck, meshgen

PLANE_1E = """numstep=100;
dt=5e-3;
interpolation=5;
nq=6;

E=5e7;
nu=0.25;
density=1730;

sd=3;
nnd=4;
if_ext=1;
EType=1;
nel=1;

nodes=
   0.00000   0.00000   0.0000
   1.0\t     0.0       0.0
   1.0       1.0       0.0
   0.0       1.0       0.0

IEN=
   1   2   3   4.


UBC_num=0;
UBC_dof=

UBC_values=


FBC_num=4;

FBC_dof=
\t3,6,9,12

FBC_values=
\t1.0, 1.0, 1.0, 1.0
"""


def test_read_deck_reference_style(tmp_path):
    f = tmp_path / "plane.txt"
    f.write_text(PLANE_1E)
    d = deck.read_deck(f)
    assert (d.numstep, d.dt, d.nq, d.E, d.nu, d.density) == (100, 5e-3, 6, 5e7, 0.25, 1730)
    assert (d.nnd, d.nel, d.if_ext, d.etype) == (4, 1, 1, 1)
    np.testing.assert_array_equal(d.ien, [[0, 1, 2, 3]])   # "4." reads as 4 (atoi)
    np.testing.assert_array_equal(d.fbc_dof, [2, 5, 8, 11])
    np.testing.assert_array_equal(d.fbc_val, [1, 1, 1, 1])
    assert d.coords.shape == (4, 3) and d.coords[2, 1] == 1.0


def test_unknown_keyword_and_bad_etype(tmp_path):
    f = tmp_path / "bad.txt"
    f.write_text("numstep=3;\nbogus=1;\n")
    with pytest.raises(deck.DeckError, match="unaccepted key word"):
        deck.read_deck(f)
    f.write_text("EType=3;\n")
    with pytest.raises(deck.DeckError, match="unsupported element type"):
        deck.read_deck(f)


def test_write_read_roundtrip(tmp_path):
    p = configs.bar(h=0.5, numstep=7)
    f = tmp_path / "bar.txt"
    deck.write_deck(f, p.coords, p.ien, E=p.E, nu=p.nu, density=p.rho, dt=p.dt,
                    numstep=p.numstep, ubc_dof=p.ubc_dof, ubc_val=p.ubc_val, fbc_dof=p.fbc_dof,
                    fbc_val=p.fbc_val)
    d = deck.read_deck(f)
    np.testing.assert_array_equal(d.coords, p.coords)
    np.testing.assert_array_equal(d.ien, p.ien)
    np.testing.assert_array_equal(d.ubc_dof, p.ubc_dof)
    np.testing.assert_array_equal(d.fbc_dof, p.fbc_dof)
    assert d.dt == p.dt and d.numstep == 7


REF_GEOM = Path("/root/reference/geometry")


@pytest.mark.skipif(not REF_GEOM.is_dir(), reason="reference decks not mounted")
def test_reference_decks_parse():
    n = 0
    for f in sorted(REF_GEOM.glob("*.txt")):
        if f.name == "Bar_550_quad.txt":
            # blank lines between IEN rows: the reference would hand atoi a NULL token
            with pytest.raises(deck.DeckError, match="fewer than 8"):
                deck.read_deck(f)
            continue
        d = deck.read_deck(f)
        if d.formulation == 3:
            assert d.coords_aux.shape == (d.nnd_aux, 3)
            assert d.ien_aux.shape == (d.nel_aux, d.ennd_aux)
        assert d.coords.shape == (d.nnd, 3)
        assert d.ien.shape == (d.nel, d.ennd)
        assert d.ien.min() >= 0 and d.ien.max() < d.nnd
        n += 1
    assert n >= 20


def _normals(m):
    X = m.coords[m.ien[:, :4]]
    n = np.cross(X[:, 1] - X[:, 0], X[:, 3] - X[:, 0])
    c = X.mean(axis=1)
    return n, c


def test_bar_mesh_counts_and_outward_normals():
    p = configs.bar()
    assert (p.nnd, p.nel, p.ndof) == (290, 288, 870)
    assert p.max_step() == 50
    n, c = _normals(meshgen.Mesh(p.coords, p.ien))
    assert np.all(np.einsum("ij,ij->i", n, c - p.coords.mean(0)) > 0)
    # closed surface: the area-weighted normals sum to zero
    assert np.abs(n.sum(0)).max() < 1e-12
    assert len(p.ubc_dof) == 75 and len(p.fbc_dof) == 25


def test_sphere_mesh_counts_inward_normals_and_pressure_sign():
    p = configs.sphere()
    assert (p.nnd, p.nel, p.ndof) == (1538, 1536, 4614)
    assert p.max_step() == 65
    n, c = _normals(meshgen.Mesh(p.coords, p.ien))
    assert np.all(np.einsum("ij,ij->i", n, c) < 0)          # into the cavity
    t = p.fbc_val.reshape(-1, 3)
    x = p.coords[p.fbc_dof.reshape(-1, 3)[:, 0] // 3]
    assert np.all(np.einsum("ij,ij->i", t, x) > 0)           # pressure pushes the wall outward
    assert np.allclose(np.linalg.norm(p.coords, axis=1), 1.0)


def test_halfspace_scenarios_are_seeded_and_uneven():
    a = configs.halfspace_scenarios(count=4, seed=4, el_range=(400, 900))
    b = configs.halfspace_scenarios(count=4, seed=4, el_range=(400, 900))
    assert [x.nel for x in a] == [x.nel for x in b]
    for s in a:
        assert 300 <= s.nel <= 1200
        assert s.if_ext == 1 and 0.2 <= s.nu <= 0.35
        cav = s.fbc_dof.reshape(-1, 3)[:, 0] // 3
        assert np.all(s.coords[cav, 2] < -0.4)               # loads on the cavity only


def test_serendipity_conversion():
    b = configs.bar(h=1.0)
    m = meshgen.to_serendipity(meshgen.Mesh(b.coords, b.ien))
    assert m.ien.shape == (b.nel, 8)
    mid = m.coords[m.ien[:, 4]]
    np.testing.assert_allclose(mid, 0.5 * (m.coords[m.ien[:, 0]] + m.coords[m.ien[:, 1]]))


def test_write_deck_round_trips_the_interface(tmp_path):
    z = np.load(Path(__file__).resolve().parent / "golden" / "layer_decks.npz")
    f = tmp_path / "region.txt"
    deck.write_deck(f, z["coords0"], z["ien0"], E=1e6, nu=0.0, density=1.0, dt=5e-3, numstep=10,
                    fbc_dof=z["fbc_dof0"], fbc_val=z["fbc_val0"], formulation=3,
                    interface=z["iface0"])
    d = deck.read_deck(f)
    np.testing.assert_array_equal(d.interface, z["iface0"])
    np.testing.assert_array_equal(d.coords, z["coords0"])
    np.testing.assert_array_equal(d.fbc_dof, z["fbc_dof0"])
