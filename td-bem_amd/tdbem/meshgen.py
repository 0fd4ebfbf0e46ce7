"""Synthetic surface meshes for the TD-BEM workloads, with the reference's orientation rules.

Normals (d x/d eta1 x d x/d eta2 with the node order 1-2-3-4) point out of the elastic domain,
as in the reference decks: outward on a bar (geometry/Bar_550_linear.txt), into the cavity for an
exterior problem (geometry/SphericalCavity_1.txt, geometry/HalfSpaceCavity_1.txt), +z on a
free surface over a lower half-space.
"""
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Mesh:
    coords: np.ndarray            # [nnd, 3] float64
    ien: np.ndarray               # [nel, 4] int32, 0-based
    etype: int = 1
    tags: dict = field(default_factory=dict)   # name -> node index array

    @property
    def nnd(self):
        return self.coords.shape[0]

    @property
    def nel(self):
        return self.ien.shape[0]


class _NodeTable:
    """Deduplicates nodes by rounded position."""

    def __init__(self, tol=1e-9):
        self.scale = 1.0 / tol
        self.index = {}
        self.coords = []

    def add(self, p):
        key = tuple(int(round(c * self.scale)) for c in p)
        i = self.index.get(key)
        if i is None:
            i = len(self.coords)
            self.index[key] = i
            self.coords.append(np.asarray(p, dtype=np.float64))
        return i


def _grid_face(table, origin, du, dv, nu_, nv_, ien, flip=False):
    ids = np.empty((nu_ + 1, nv_ + 1), dtype=np.int64)
    for a in range(nu_ + 1):
        for b in range(nv_ + 1):
            ids[a, b] = table.add(origin + du * (a / nu_) + dv * (b / nv_))
    for a in range(nu_):
        for b in range(nv_):
            quad = [ids[a, b], ids[a + 1, b], ids[a + 1, b + 1], ids[a, b + 1]]
            if flip:
                quad = [quad[0], quad[3], quad[2], quad[1]]
            ien.append(quad)


def box_surface(lx=1.0, ly=1.0, lz=4.0, h=0.25):
    """Closed box [0,lx]x[0,ly]x[0,lz] of bilinear quads of size h, outward normals."""
    nx, ny, nz = (int(round(v / h)) for v in (lx, ly, lz))
    t = _NodeTable()
    ien = []
    o = np.zeros(3)
    ex, ey, ez = np.array([lx, 0, 0.]), np.array([0, ly, 0.]), np.array([0, 0, lz])
    # z = 0 (normal -z): ey x ex = -z
    _grid_face(t, o, ey, ex, ny, nx, ien)
    # z = lz (normal +z)
    _grid_face(t, o + ez, ex, ey, nx, ny, ien)
    # x = 0 (normal -x): ez x ey = -x
    _grid_face(t, o, ez, ey, nz, ny, ien)
    # x = lx (normal +x)
    _grid_face(t, o + ex, ey, ez, ny, nz, ien)
    # y = 0 (normal -y): ex x ez = -y
    _grid_face(t, o, ex, ez, nx, nz, ien)
    # y = ly (normal +y)
    _grid_face(t, o + ey, ez, ex, nz, nx, ien)
    coords = np.array(t.coords)
    m = Mesh(coords, np.array(ien, dtype=np.int32))
    m.tags["z0"] = np.where(np.abs(coords[:, 2]) < 1e-12)[0]
    m.tags["ztop"] = np.where(np.abs(coords[:, 2] - lz) < 1e-12)[0]
    return m


def cube_sphere(n=16, radius=1.0, center=(0.0, 0.0, 0.0), inward=True):
    """Gnomonic (equiangular) cube-sphere with 6*n*n quads and 6*n*n+2 nodes."""
    t = _NodeTable(tol=1e-9 * max(radius, 1.0))
    ien = []
    c = np.asarray(center, dtype=np.float64)
    ang = np.tan(np.linspace(-np.pi / 4, np.pi / 4, n + 1))
    # each face: (normal axis, sign) with a right-handed (u, v) so that u x v = outward
    faces = [
        (np.array([1., 0, 0]), np.array([0, 1., 0]), np.array([0, 0, 1.])),
        (np.array([-1., 0, 0]), np.array([0, 0, 1.]), np.array([0, 1., 0])),
        (np.array([0, 1., 0]), np.array([0, 0, 1.]), np.array([1., 0, 0])),
        (np.array([0, -1., 0]), np.array([1., 0, 0]), np.array([0, 0, 1.])),
        (np.array([0, 0, 1.]), np.array([1., 0, 0]), np.array([0, 1., 0])),
        (np.array([0, 0, -1.]), np.array([0, 1., 0]), np.array([1., 0, 0])),
    ]
    for nrm, u, v in faces:
        ids = np.empty((n + 1, n + 1), dtype=np.int64)
        for a in range(n + 1):
            for b in range(n + 1):
                p = nrm + ang[a] * u + ang[b] * v
                p = c + radius * p / np.linalg.norm(p)
                ids[a, b] = t.add(p)
        for a in range(n):
            for b in range(n):
                quad = [ids[a, b], ids[a + 1, b], ids[a + 1, b + 1], ids[a, b + 1]]
                if inward:
                    quad = [quad[0], quad[3], quad[2], quad[1]]
                ien.append(quad)
    return Mesh(np.array(t.coords), np.array(ien, dtype=np.int32))


def plane(size=8.0, n=16, z=0.0):
    """Square [-size/2, size/2]^2 at height z with +z normals (free surface over z < z0)."""
    t = _NodeTable()
    ien = []
    o = np.array([-size / 2, -size / 2, z])
    _grid_face(t, o, np.array([size, 0, 0.]), np.array([0, size, 0.]), n, n, ien)
    return Mesh(np.array(t.coords), np.array(ien, dtype=np.int32))


def merge(*meshes):
    coords = []
    ien = []
    tags = {}
    off = 0
    for i, m in enumerate(meshes):
        coords.append(m.coords)
        ien.append(m.ien + off)
        tags[f"part{i}"] = np.arange(off, off + m.nnd)
        off += m.nnd
    return Mesh(np.vstack(coords), np.vstack(ien).astype(np.int32), tags=tags)


def halfspace_cavity(n_elements, radius, depth, plane_size=8.0):
    """Truncated free-surface plane plus a spherical cavity centred at depth `depth`, with the
    element count split so both parts have about the same element size."""
    area_plane = plane_size ** 2
    area_sphere = 4 * np.pi * radius ** 2
    h = np.sqrt((area_plane + area_sphere) / n_elements)
    n_p = max(2, int(round(plane_size / h)))
    n_s = max(2, int(round(np.sqrt(area_sphere / 6) / h)))
    m = merge(plane(plane_size, n_p), cube_sphere(n_s, radius, (0.0, 0.0, -depth), inward=True))
    m.tags["plane"] = m.tags.pop("part0")
    m.tags["cavity"] = m.tags.pop("part1")
    m.tags["center"] = np.array([0.0, 0.0, -depth])
    return m


def element_size(mesh):
    """Mean square root of element area (the h of beta = c_p dt / h)."""
    X = mesh.coords[mesh.ien[:, :4]]
    a = np.cross(X[:, 1] - X[:, 0], X[:, 3] - X[:, 0])
    b = np.cross(X[:, 3] - X[:, 2], X[:, 1] - X[:, 2])
    area = 0.5 * (np.linalg.norm(a, axis=1) + np.linalg.norm(b, axis=1))
    return float(np.sqrt(area.mean()))


def to_serendipity(mesh):
    """8-node quads from a bilinear mesh: midside nodes 5-8 on edges 1-2, 2-3, 3-4, 4-1
    (the node order of ShapeFunctionT type 2, src/ShapeFunctionT.cpp:304-311)."""
    t = _NodeTable()
    for p in mesh.coords:
        t.add(p)
    ien = []
    for q in mesh.ien:
        mids = []
        for a, b in ((0, 1), (1, 2), (2, 3), (3, 0)):
            mids.append(t.add(0.5 * (mesh.coords[q[a]] + mesh.coords[q[b]])))
        ien.append(list(q[:4]) + mids)
    return Mesh(np.array(t.coords), np.array(ien, dtype=np.int32), etype=2, tags=dict(mesh.tags))
