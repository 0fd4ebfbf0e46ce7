"""ctypes loader for the CPU oracle (oracle/liboracle.so) and, when built, the reference harness
(oracle/_ref/libtdbem_ref.so).  Test infrastructure: only tests/, smoke() and bench.py's
cpu_baseline leg import this module."""
import ctypes as C
import os
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_SO = ROOT / "oracle" / "liboracle.so"
REF_SO = ROOT / "oracle" / "_ref" / "libtdbem_ref.so"

P = np.ctypeslib.ndpointer
D = C.c_double
I = C.c_int
dp = C.POINTER(C.c_double)


class Material(C.Structure):
    _fields_ = [("E", D), ("nu", D), ("rho", D), ("mu", D), ("cp", D), ("cs", D)]


class MarchCfg(C.Structure):
    _fields_ = [("N", I), ("K", I), ("numstep", I), ("dt", D), ("a1", D), ("a2", D),
                ("n_ubc", I), ("ubc_dof", C.POINTER(C.c_int)), ("ubc_val", dp),
                ("n_fbc", I), ("fbc_dof", C.POINTER(C.c_int)), ("fbc_val", dp),
                ("amp", D), ("t_rise", D)]


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _ptr(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        if not ORACLE_SO.exists():
            raise RuntimeError(f"{ORACLE_SO} is not built (run __graft_entry__.build())")
        lib = C.CDLL(str(ORACLE_SO))
        lib.or_load_factor.restype = D
        lib.or_mrt_load_factor.restype = D
        _oracle = lib
    return _oracle


def ref():
    """The reference's compiled classes, or None when oracle/_ref was not built."""
    global _ref
    if _ref is None and REF_SO.exists():
        _ref = C.CDLL(str(REF_SO))
    return _ref


def material(E, nu, rho):
    m = Material()
    oracle().or_material_set(C.byref(m), D(E), D(nu), D(rho))
    return m


def element(kernel, etype, nq, E, nu, rho, coll, X, t, dt, step_index, singular, use_ref=False):
    ennd = 4 if etype == 1 else 8
    out = [np.zeros(3 * 3 * ennd) for _ in range(4)] + [np.zeros(9), np.zeros(9)]
    coll = _f64(coll)
    X = _f64(X)
    if use_ref:
        rc = ref().ref_element(I(kernel), I(etype), I(nq), D(E), D(nu), D(rho), _ptr(coll),
                               _ptr(X), D(t), D(dt), I(singular), *[_ptr(o) for o in out])
    else:
        m = material(E, nu, rho)
        rc = oracle().or_element(I(kernel), I(etype), I(nq), C.byref(m), _ptr(coll), _ptr(X),
                                 D(t), D(dt), I(step_index), I(singular), *[_ptr(o) for o in out])
    if rc:
        raise RuntimeError(f"element failed rc={rc}")
    return [o.reshape(3, -1) if i < 4 else o.reshape(3, 3) for i, o in enumerate(out)]


def green(kernel, E, nu, rho, direction, normal, t, dt, use_ref=False):
    """Returns DU1, DU2, DT1, DT2, K1, K2 (each 3x3) at one point."""
    d = _f64(direction)
    n = _f64(normal)
    if use_ref:
        out = np.zeros(54)
        if kernel == 6:
            pt, g = np.zeros(18), np.zeros(2)
            ref().ref_green_ramp(D(E), D(nu), D(rho), _ptr(d), _ptr(n), D(t), D(dt), _ptr(out),
                                 _ptr(pt), _ptr(g))
        else:
            ref().ref_green(I(kernel), D(E), D(nu), D(rho), _ptr(d), _ptr(n), D(t), D(dt),
                            _ptr(out))
        return [out[9 * i:9 * i + 9].reshape(3, 3) for i in range(6)]
    m = material(E, nu, rho)
    o = [np.zeros(9) for _ in range(6)]
    lib = oracle()
    if kernel == 4:
        lib.or_stokes_step_linear(C.byref(m), _ptr(d), _ptr(n), D(t), D(dt), *[_ptr(x) for x in o[:4]])
        lib.or_kelvin_linear(C.byref(m), _ptr(d), _ptr(n), D(dt), _ptr(o[4]))
        o[5][:] = o[4]
    elif kernel == 2:
        lib.or_stokes_heaviside_linear(C.byref(m), _ptr(d), _ptr(n), D(t), D(dt), *[_ptr(x) for x in o[:4]])
        lib.or_kelvin_linear(C.byref(m), _ptr(d), _ptr(n), D(dt), _ptr(o[4]))
        o[5][:] = o[4]
    elif kernel == 5:
        lib.or_stokes_bspline_linear(C.byref(m), _ptr(d), _ptr(n), D(t), D(dt), *[_ptr(x) for x in o[:4]])
        lib.or_kelvin_bspline_linear(C.byref(m), _ptr(d), _ptr(n), D(t), D(dt), _ptr(o[4]), _ptr(o[5]))
    elif kernel == 6:
        lib.or_stokes_ramp_linear(C.byref(m), _ptr(d), _ptr(n), D(t), D(dt), *[_ptr(x) for x in o[:4]])
        lib.or_kelvin_ramp_linear(C.byref(m), _ptr(d), _ptr(n), D(t), D(dt), _ptr(o[4]), _ptr(o[5]))
    else:
        raise ValueError(f"kernel {kernel}")
    return [x.reshape(3, 3) for x in o]


def ramp_point(E, nu, rho, direction, normal, t, dt, use_ref=False):
    """Stokes_RampedHeaviside at time t (slope 2/dt): U, T (3x3 each), and the
    Int_RampedHeaviside_Linear weights (g1, g2) at the same t."""
    d = _f64(direction)
    n = _f64(normal)
    if use_ref:
        out, pt, g = np.zeros(54), np.zeros(18), np.zeros(2)
        ref().ref_green_ramp(D(E), D(nu), D(rho), _ptr(d), _ptr(n), D(t), D(dt), _ptr(out),
                             _ptr(pt), _ptr(g))
        return pt[:9].reshape(3, 3), pt[9:].reshape(3, 3), g
    m = material(E, nu, rho)
    U, T, g = np.zeros(9), np.zeros(9), np.zeros(2)
    oracle().or_stokes_ramp_point(C.byref(m), _ptr(d), _ptr(n), D(t), D(dt), _ptr(U), _ptr(T))
    oracle().or_ramp_weights(D(t), D(dt), _ptr(g))
    return U.reshape(3, 3), T.reshape(3, 3), g


def shape_eval(etype, X, e1, e2, use_ref=False):
    ennd = 4 if etype == 1 else 8
    X = _f64(X)
    if use_ref:
        out = np.zeros(ennd + 7)
        rc = ref().ref_shape_eval(I(etype), _ptr(X), D(e1), D(e2), _ptr(out))
        if rc:
            raise RuntimeError("shape eval failed")
        return out[:ennd], out[ennd:ennd + 3], out[ennd + 3:ennd + 6], out[ennd + 6]
    N = np.zeros(ennd); x = np.zeros(3); n = np.zeros(3); J = D()
    rc = oracle().or_shape_eval(I(etype), _ptr(X), D(e1), D(e2), _ptr(N), _ptr(x), _ptr(n), C.byref(J))
    if rc:
        raise RuntimeError("shape eval failed")
    return N, x, n, J.value


def assemble_step(kernel, etype, nq, E, nu, rho, coords, ien, if_ext, lower, upper, t, dt,
                  step_index, use_ref=False, nthreads=0):
    coords = _f64(coords)
    ien = _i32(ien)
    nnd = coords.shape[0]
    nel = ien.shape[0]
    nrow = 3 * (upper - lower + 1)
    out = [np.zeros((nrow, 3 * nnd)) for _ in range(4)]
    if use_ref:
        rc = ref().ref_assemble_step(I(kernel), I(etype), I(nq), D(E), D(nu), D(rho), I(nnd),
                                     _ptr(coords), I(nel), _ptr(ien, C.c_int), I(if_ext),
                                     I(lower), I(upper), D(t), D(dt), *[_ptr(o) for o in out])
    else:
        m = material(E, nu, rho)
        rc = oracle().or_assemble_step(I(kernel), I(etype), I(nq), C.byref(m), I(nnd),
                                       _ptr(coords), I(nel), _ptr(ien, C.c_int), I(if_ext),
                                       I(lower), I(upper), D(t), D(dt), I(step_index),
                                       *[_ptr(o) for o in out], I(nthreads))
    if rc:
        raise RuntimeError(f"assemble_step failed rc={rc}")
    return out


def max_step(E, nu, rho, coords, dt):
    m = material(E, nu, rho)
    coords = _f64(coords)
    return oracle().or_max_step(C.byref(m), I(coords.shape[0]), _ptr(coords), D(dt))


def build_lags(kernel, etype, nq, E, nu, rho, coords, ien, if_ext, dt, numstep, K, nthreads=0):
    coords = _f64(coords)
    ien = _i32(ien)
    N = 3 * coords.shape[0]
    MG = np.zeros((K + 1, N, N))
    MH = np.zeros((K + 1, N, N))
    pg = (dp * (K + 1))(*[_ptr(MG[k]) for k in range(K + 1)])
    ph = (dp * (K + 1))(*[_ptr(MH[k]) for k in range(K + 1)])
    m = material(E, nu, rho)
    rc = oracle().or_build_lags(I(kernel), I(etype), I(nq), C.byref(m), I(coords.shape[0]),
                                _ptr(coords), I(ien.shape[0]), _ptr(ien, C.c_int), I(if_ext),
                                D(dt), I(numstep), I(K), pg, ph, I(nthreads))
    if rc:
        raise RuntimeError(f"build_lags failed rc={rc}")
    return MG, MH


def march(MG, MH, numstep, dt, ubc_dof, ubc_val, fbc_dof, fbc_val, a1=0.25, a2=0.25, amp=1.0,
          t_rise=0.0, fused=False):
    K = MG.shape[0] - 1
    N = MG.shape[1]
    ubc_dof = _i32(ubc_dof); ubc_val = _f64(ubc_val)
    fbc_dof = _i32(fbc_dof); fbc_val = _f64(fbc_val)
    cfg = MarchCfg(N, K, numstep, dt, a1, a2, len(ubc_dof), _ptr(ubc_dof, C.c_int), _ptr(ubc_val),
                   len(fbc_dof), _ptr(fbc_dof, C.c_int), _ptr(fbc_val), amp, t_rise)
    MG = np.ascontiguousarray(MG); MH = np.ascontiguousarray(MH)
    pg = (dp * (K + 1))(*[_ptr(MG[k]) for k in range(K + 1)])
    ph = (dp * (K + 1))(*[_ptr(MH[k]) for k in range(K + 1)])
    U = np.zeros((numstep, N)); T = np.zeros((numstep, N))
    fn = oracle().or_march_fused if fused else oracle().or_march_literal
    rc = fn(C.byref(cfg), pg, ph, _ptr(U), _ptr(T))
    if rc == -1:
        raise ValueError("Alpha is out of range")
    if rc:
        raise RuntimeError(f"march failed rc={rc}")
    return U, T


class MrtCfg(C.Structure):
    _fields_ = [("nnd", I * 2), ("coords", dp * 2), ("n_iface", I), ("iface", C.POINTER(C.c_int) * 2),
                ("n_fbc", I * 2), ("fbc_dof", C.POINTER(C.c_int) * 2), ("fbc_val", dp * 2),
                ("K", I), ("numstep", I), ("scheme", I), ("load_kind", I),
                ("dt", D), ("amp", D), ("t_rise", D)]


def mrt_march(regions, lags, numstep, dt, K, scheme=0, load_kind=0, amp=1.0, t_rise=0.0):
    """SolverMRT restated (or_mrt_march): regions = two dicts with coords, iface (0-based node
    ids), fbc_dof, fbc_val; lags = [(MG1, MH1), (MG2, MH2)] with K + 1 lags each.  Returns the
    global unknowns X [numstep+1][3 (nnd1 + nnd2)] and the region displacements [2][numstep+1][3 nnd]."""
    keep = []

    def arr(a, t):
        a = np.ascontiguousarray(a, dtype=t)
        keep.append(a)
        return a
    cfg = MrtCfg()
    for r, reg in enumerate(regions):
        co = arr(reg["coords"], np.float64)
        cfg.nnd[r] = co.shape[0]
        cfg.coords[r] = _ptr(co)
        cfg.iface[r] = _ptr(arr(reg["iface"], np.int32), C.c_int)
        fd = arr(reg["fbc_dof"], np.int32)
        cfg.n_fbc[r] = len(fd)
        cfg.fbc_dof[r] = _ptr(fd, C.c_int)
        cfg.fbc_val[r] = _ptr(arr(reg["fbc_val"], np.float64))
    cfg.n_iface = len(regions[0]["iface"])
    cfg.K, cfg.numstep, cfg.scheme, cfg.load_kind = K, numstep, scheme, load_kind
    cfg.dt, cfg.amp, cfg.t_rise = dt, amp, t_rise
    ptrs = []
    for MG, MH in lags:
        MG = arr(MG, np.float64); MH = arr(MH, np.float64)
        ptrs.append(((dp * (K + 1))(*[_ptr(MG[k]) for k in range(K + 1)]),
                     (dp * (K + 1))(*[_ptr(MH[k]) for k in range(K + 1)])))
    NH = 3 * (cfg.nnd[0] + cfg.nnd[1])
    X = np.zeros((numstep + 1, NH))
    rc = oracle().or_mrt_march(C.byref(cfg), ptrs[0][0], ptrs[0][1], ptrs[1][0], ptrs[1][1], _ptr(X))
    if rc == -1:
        raise ValueError("Interface mismatch, interface nodes must match one-to-one")
    if rc:
        raise RuntimeError(f"mrt march failed rc={rc}")
    U = []
    for r in range(2):
        u = np.zeros((numstep + 1, 3 * cfg.nnd[r]))
        for s in range(numstep + 1):
            oracle().or_mrt_extract(C.byref(cfg), I(r), _ptr(X[s]), _ptr(u[s]))
        U.append(u)
    return X, U
