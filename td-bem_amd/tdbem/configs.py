"""The five workloads: geometry, material, boundary conditions and time step.

beta = c_p dt / h with h = sqrt(mean element area) (0.25 on the bar; 0.0905 on the 6x16x16
sphere, which gives MaxStep = 65 with the reference's formula, src/SolverT.cpp:211-250).
"""
from dataclasses import dataclass, field

import numpy as np

from . import meshgen


@dataclass
class Problem:
    name: str
    coords: np.ndarray
    ien: np.ndarray
    E: float
    nu: float
    rho: float
    dt: float
    numstep: int
    if_ext: int
    ubc_dof: np.ndarray
    ubc_val: np.ndarray
    fbc_dof: np.ndarray
    fbc_val: np.ndarray
    nq: int = 6
    etype: int = 1
    kernel: int = 4          # 4 = Step (TDElementStepLinearT, the reference's main path)
    a1: float = 0.25
    a2: float = 0.25
    amp: float = 1.0         # load history f(t) = amp * min(1, t / t_rise)
    t_rise: float = 0.0
    meta: dict = field(default_factory=dict)

    @property
    def nnd(self):
        return self.coords.shape[0]

    @property
    def nel(self):
        return self.ien.shape[0]

    @property
    def ndof(self):
        return 3 * self.nnd

    def wave_speeds(self):
        mu = self.E / (2 * (1 + self.nu))
        lam = self.E * self.nu / ((1 + self.nu) * (1 - 2 * self.nu))
        return np.sqrt((lam + 2 * mu) / self.rho), np.sqrt(mu / self.rho)

    def max_step(self):
        """SolverT::ComputeMaxStep (src/SolverT.cpp:211-250)."""
        X = self.coords
        d2 = 0.0
        for i in range(0, len(X), 512):
            blk = X[i:i + 512]
            d2 = max(d2, float(((blk[:, None, :] - X[None, :, :]) ** 2).sum(-1).max()))
        mu = self.E / (2.0 * (1.0 + self.nu))
        cs = np.sqrt(mu / self.rho)
        return int(np.ceil(np.sqrt(d2) / (cs * self.dt))) + 1


def _dt(E, nu, rho, h, beta):
    mu = E / (2 * (1 + nu))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    cp = np.sqrt((lam + 2 * mu) / rho)
    return beta * h / cp


def bar(h=0.25, length=4.0, numstep=100, beta=0.6, E=1.0, nu=0.25, rho=1.0):
    """Prismatic bar 1x1xL: u = 0 on z = 0, unit Heaviside traction t_z on z = L."""
    m = meshgen.box_surface(1.0, 1.0, length, h)
    ubc = np.concatenate([3 * m.tags["z0"] + d for d in range(3)]).astype(np.int32)
    ubc.sort()
    fbc = (3 * m.tags["ztop"] + 2).astype(np.int32)
    fbc.sort()
    return Problem("bar", m.coords, m.ien, E, nu, rho, _dt(E, nu, rho, h, beta), numstep, 0,
                   ubc, np.zeros(len(ubc)), fbc, np.ones(len(fbc)), meta={"h": h})


def split_bar(h=0.5, length=4.0, numstep=40, beta=0.6, E=1.0, nu=0.25, rho=1.0):
    """The free 1x1xL bar cut at z = L/2 into two regions for the two-region solver: region 1
    [0,1]^2 x [0,L/2], region 2 above it, interface z = L/2 (region 1's top face, region 2's
    bottom face), unit traction t_z on z = L.  Returns (region1, region2, iface1, iface2)."""
    dt = _dt(E, nu, rho, h, beta)
    m1 = meshgen.box_surface(1.0, 1.0, length / 2, h)
    m2 = meshgen.box_surface(1.0, 1.0, length / 2, h)
    c2 = m2.coords.copy()
    c2[:, 2] += length / 2
    top = m2.tags["ztop"]
    fbc = np.sort(3 * top + 2).astype(np.int32)
    none_i, none_v = np.zeros(0, np.int32), np.zeros(0)
    r1 = Problem("split_bar_lower", m1.coords, m1.ien, E, nu, rho, dt, numstep, 0, none_i, none_v,
                 none_i, none_v, meta={"h": h})
    r2 = Problem("split_bar_upper", c2, m2.ien, E, nu, rho, dt, numstep, 0, none_i, none_v, fbc,
                 np.ones(len(fbc)), meta={"h": h})
    return r1, r2, m1.tags["ztop"].astype(np.int32), m2.tags["z0"].astype(np.int32)


def _pressure_fbc(coords, center, p, nodes=None):
    """Traction t = p * (x - c)/|x - c| on every dof of `nodes` (a pressure p in a cavity pushes
    the wall outward; the element normals point into the cavity)."""
    nodes = np.arange(len(coords)) if nodes is None else np.asarray(nodes)
    r = coords[nodes] - np.asarray(center)[None, :]
    r /= np.linalg.norm(r, axis=1)[:, None]
    dof = (3 * nodes[:, None] + np.arange(3)[None, :]).reshape(-1).astype(np.int32)
    val = (p * r).reshape(-1)
    return dof, val


def sphere(n=16, numstep=600, beta=0.6, E=1.0, nu=0.25, rho=1.0, p=1.0):
    """Spherical cavity radius 1 in full space, Heaviside internal pressure p."""
    m = meshgen.cube_sphere(n, 1.0, inward=True)
    h = meshgen.element_size(m)
    fdof, fval = _pressure_fbc(m.coords, (0, 0, 0), p)
    return Problem("sphere", m.coords, m.ien, E, nu, rho, _dt(E, nu, rho, h, beta), numstep, 1,
                   np.zeros(0, np.int32), np.zeros(0), fdof, fval, meta={"h": h, "p": p})


def rhs_set(count=32, dt=1.0, seed=3):
    """Per-RHS load amplitude U[0.5, 2] and rise time U[0, 10 dt] (0 = Heaviside)."""
    rng = np.random.default_rng(seed)
    amp = rng.uniform(0.5, 2.0, count)
    rise = rng.uniform(0.0, 10.0 * dt, count)
    rise[rng.uniform(size=count) < 0.5] = 0.0
    return amp, rise


def halfspace_scenarios(count=64, seed=4, el_range=(400, 4000), step_range=(100, 800),
                        beta=0.6, p=1.0):
    """Half-space cavity problems of uneven size: free plane 8x8 + spherical cavity."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(count):
        nel = int(rng.integers(el_range[0], el_range[1] + 1))
        radius = float(rng.uniform(0.5, 1.5))
        depth = float(rng.uniform(2.0, 4.0))
        nu = float(rng.uniform(0.2, 0.35))
        steps = int(rng.integers(step_range[0], step_range[1] + 1))
        m = meshgen.halfspace_cavity(nel, radius, depth)
        h = meshgen.element_size(m)
        fdof, fval = _pressure_fbc(m.coords, m.tags["center"], p, m.tags["cavity"])
        out.append(Problem(f"halfspace{s}", m.coords, m.ien, 1.0, nu, 1.0,
                           _dt(1.0, nu, 1.0, h, beta), steps, 1, np.zeros(0, np.int32),
                           np.zeros(0), fdof, fval,
                           meta={"h": h, "radius": radius, "depth": depth}))
    return out


def serving_mix(count=16, seed=5, el_range=(300, 6000), step_range=(100, 1000)):
    """Solve requests for the serving loop from the bar and sphere geometries: element counts
    log-uniform in el_range (many small solves, few large ones), each taken to the nearest mesh of
    its geometry inside the range (bar: 18 n^2 quads, n = 5..18 for the default range; sphere:
    6 n^2, n = 8..31), step counts U[step_range]."""
    rng = np.random.default_rng(seed)
    lo, hi = el_range
    out = []
    for _ in range(count):
        kind = "bar" if rng.uniform() < 0.5 else "sphere"
        target = float(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        steps = int(rng.integers(step_range[0], step_range[1] + 1))
        per = 18 if kind == "bar" else 6
        sizes = [n for n in range(2, 64) if lo <= per * n * n <= hi]
        n = min(sizes, key=lambda m: abs(per * m * m - target))
        p = bar(h=1.0 / n, numstep=steps) if kind == "bar" else sphere(n=n, numstep=steps)
        out.append(p)
    return out
