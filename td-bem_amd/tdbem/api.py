"""ctypes binding of the C ABI in include/tdbem.h (libtdbem.so built from td-bem_amd/csrc).

This is the Python face of the same entry points the C++ host (td-bem_amd/host) calls.  Loading
fails loudly when the HIP library is missing: there is no CPU fallback.
"""
import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parents[1]
# TDBEM_LIBRARY selects another build of the same library (same C ABI)
LIB_PATH = Path(os.environ.get("TDBEM_LIBRARY", str(PKG / "libtdbem.so")))

ERR_ARG, ERR_HIP, ERR_STATE, ERR_NUMERIC = -1, -2, -3, -4
KERNEL_HEAVISIDE = 2
KERNEL_STEP = 4
KERNEL_BSPLINE = 5
KERNEL_RAMP = 6
PRECISION_FP64 = 0
PRECISION_MIXED = 1
MRT_DIRECT, MRT_AVERAGING = 0, 1
MRT_LOAD_PULSE, MRT_LOAD_RAMP = 0, 1
PROFILE_ALL, PROFILE_CONV = 1, 2

_lib = None


class TdbemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"tdbem error {code}: {msg}")
        self.code = code
        self.msg = msg


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(str(LIB_PATH))
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        vp = C.c_void_p
        L.tdbem_create.argtypes = [C.c_int, dp, C.c_int, ip, C.c_int, C.c_int, C.c_double,
                                   C.c_double, C.c_double, C.c_int, C.c_double, C.c_int, C.c_int,
                                   C.POINTER(vp)]
        L.tdbem_destroy.argtypes = [vp]
        L.tdbem_last_error.argtypes = [vp]
        L.tdbem_last_error.restype = C.c_char_p
        L.tdbem_max_step.argtypes = [vp]
        L.tdbem_assemble_step.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp, dp, dp, dp]
        L.tdbem_assemble_step_device.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp]
        L.tdbem_build_lag_store.argtypes = [vp]
        L.tdbem_lag_store_bytes.argtypes = [vp]
        L.tdbem_lag_store_bytes.restype = C.c_int64
        L.tdbem_get_lag.argtypes = [vp, C.c_int, dp, dp]
        L.tdbem_get_lag_rows.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp, dp]
        L.tdbem_update_gh_linear.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp, dp, dp, dp]
        L.tdbem_predict_bytes.argtypes = [vp, C.c_int, C.POINTER(C.c_int64),
                                          C.POINTER(C.c_int64)]
        L.tdbem_device_memory.argtypes = [C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.tdbem_march_init.argtypes = [vp, C.c_double, C.c_double, C.c_int, ip, dp, C.c_int, ip,
                                       dp, C.c_int, dp, dp]
        L.tdbem_march_run.argtypes = [vp, C.c_int]
        L.tdbem_march_step_index.argtypes = [vp]
        L.tdbem_get_history.argtypes = [vp, C.c_int, dp, dp]
        L.tdbem_synchronize.argtypes = [vp]
        L.tdbem_march_enqueue.argtypes = [vp, C.c_int]
        L.tdbem_query.argtypes = [vp]
        L.tdbem_set_precision.argtypes = [vp, C.c_int]
        L.tdbem_mfma_selftest.argtypes = [C.POINTER(C.c_float)] * 3
        L.tdbem_stream_copy.argtypes = [C.c_size_t, C.c_int, dp]
        L.tdbem_stream_read.argtypes = [C.c_size_t, C.c_int, dp]
        L.tdbem_mfma_f64_selftest.argtypes = [dp, dp, dp]
        L.tdbem_dense_solve.argtypes = [C.c_int, dp, C.c_int, dp, C.c_int]
        L.tdbem_batch_create.argtypes = [C.POINTER(vp), C.c_int, C.POINTER(vp)]
        L.tdbem_batch_add.argtypes = [vp, vp]
        L.tdbem_batch_remove.argtypes = [vp, vp]
        L.tdbem_batch_size.argtypes = [vp]
        L.tdbem_batch_tune.argtypes = [vp, dp]
        L.tdbem_batch_set_params.argtypes = [vp, C.c_int, C.c_int]
        L.tdbem_batch_run.argtypes = [vp, C.c_int, C.c_int, ip]
        L.tdbem_batch_destroy.argtypes = [vp]
        L.tdbem_get_state.argtypes = [vp, C.c_int, C.c_int, dp, dp]
        L.tdbem_set_profiling.argtypes = [vp, C.c_int]
        L.tdbem_reset_timings.argtypes = [vp]
        L.tdbem_get_timings.argtypes = [vp, dp]
        L.tdbem_set_max_step.argtypes = [vp, C.c_int]
        L.tdbem_mrt_create.argtypes = [vp, vp, C.c_int, C.c_int, ip, ip, C.POINTER(vp)]
        L.tdbem_mrt_init.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, ip,
                                     dp, C.c_int, ip, dp]
        L.tdbem_mrt_run.argtypes = [vp, C.c_int]
        L.tdbem_mrt_step_index.argtypes = [vp]
        L.tdbem_mrt_num_unknowns.argtypes = [vp]
        L.tdbem_mrt_get_state.argtypes = [vp, C.c_int, dp, dp, dp]
        L.tdbem_mrt_get_timings.argtypes = [vp, dp, dp]
        L.tdbem_mrt_destroy.argtypes = [vp]
        L.tdbem_element_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                           C.c_double, C.POINTER(vp)]
        L.tdbem_element_destroy.argtypes = [vp]
        L.tdbem_element_last_error.argtypes = [vp]
        L.tdbem_element_last_error.restype = C.c_char_p
        L.tdbem_element_form.argtypes = [vp, C.c_int64, dp, dp, ip, C.c_double, C.c_double,
                                         dp, dp, dp, dp, dp, dp]
        L.tdbem_element_form_device.argtypes = [vp, C.c_int64, vp, vp, vp, C.c_double,
                                                C.c_double, vp]
        L.tdbem_element_synchronize.argtypes = [vp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class Context:
    """One problem on the GPU: mesh, material, time step, lag store and march state."""

    def __init__(self, coords, ien, *, E, nu, rho, dt, numstep, nq=6, etype=1, if_ext=0,
                 kernel=KERNEL_STEP):
        L = lib()
        self.coords = np.ascontiguousarray(coords, dtype=np.float64)
        self.ien = np.ascontiguousarray(ien, dtype=np.int32)
        self.nnd = self.coords.shape[0]
        self.N = 3 * self.nnd
        self.numstep = int(numstep)
        h = C.c_void_p()
        rc = L.tdbem_create(self.nnd, _d(self.coords), self.ien.shape[0], _i(self.ien), etype,
                            nq, E, nu, rho, if_ext, dt, numstep, kernel, C.byref(h))
        if rc:
            raise TdbemError(rc, L.tdbem_last_error(None).decode())
        self._h = h

    @classmethod
    def from_problem(cls, p, numstep=None):
        return cls(p.coords, p.ien, E=p.E, nu=p.nu, rho=p.rho, dt=p.dt,
                   numstep=p.numstep if numstep is None else numstep, nq=p.nq, etype=p.etype,
                   if_ext=p.if_ext, kernel=p.kernel)

    def _check(self, rc):
        if rc:
            raise TdbemError(rc, lib().tdbem_last_error(self._h).decode())

    def close(self):
        """Frees the context; raises TdbemError (ERR_STATE) while it is a member of a batch."""
        if getattr(self, "_h", None):
            rc = lib().tdbem_destroy(self._h)
            if rc:
                raise TdbemError(rc, lib().tdbem_last_error(self._h).decode())
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def max_step(self):
        return lib().tdbem_max_step(self._h)

    def set_max_step(self, K):
        """SetMaxStep: keep lags 0..K (a change drops a built lag store)."""
        self._check(lib().tdbem_set_max_step(self._h, int(K)))

    def assemble_step(self, step, lower=0, upper=None):
        upper = self.nnd - 1 if upper is None else upper
        shape = (3 * (upper - lower + 1), self.N)
        out = [np.zeros(shape) for _ in range(4)]
        self._check(lib().tdbem_assemble_step(self._h, step, lower, upper, *[_d(o) for o in out]))
        return out

    def build_lag_store(self):
        self._check(lib().tdbem_build_lag_store(self._h))

    def lag_store_bytes(self):
        v = lib().tdbem_lag_store_bytes(self._h)
        if v < 0:
            self._check(int(v))
        return int(v)

    def update_gh_linear(self, step, G0, G1, H0, H1, lower=0, upper=None):
        """UpdateGHLinear at fCurrStep = step with caller-supplied row blocks: G0/H0 the phi_1
        part, G1/H1 the phi_2 part of the step, rows of collocation nodes lower..upper."""
        upper = self.nnd - 1 if upper is None else upper
        blocks = [np.ascontiguousarray(b, dtype=np.float64) for b in (G0, G1, H0, H1)]
        for b in blocks:
            if b.size != 3 * (upper - lower + 1) * self.N:
                raise ValueError("block size does not match the row range")
        self._check(lib().tdbem_update_gh_linear(self._h, int(step), int(lower), int(upper),
                                                 *[_d(b) for b in blocks]))

    def predict_bytes(self, nrhs):
        """(resident, peak) device bytes of a solve at nrhs right-hand sides in the current
        precision, from the store planner's exact counts (tdbem_predict_bytes)."""
        r, p = C.c_int64(0), C.c_int64(0)
        self._check(lib().tdbem_predict_bytes(self._h, int(nrhs), C.byref(r), C.byref(p)))
        return r.value, p.value

    def get_lag(self, k):
        G = np.zeros((self.N, self.N))
        H = np.zeros((self.N, self.N))
        self._check(lib().tdbem_get_lag(self._h, k, _d(G), _d(H)))
        return G, H

    def get_lag_rows(self, k, lower, upper):
        """Rows of collocation nodes lower..upper of lag k: G, H [3 (upper - lower + 1)][N]."""
        G = np.zeros((3 * (upper - lower + 1), self.N))
        H = np.zeros_like(G)
        self._check(lib().tdbem_get_lag_rows(self._h, int(k), int(lower), int(upper), _d(G),
                                             _d(H)))
        return G, H

    def march_init(self, a1=0.25, a2=0.25, ubc_dof=(), ubc_val=(), fbc_dof=(), fbc_val=(),
                   amp=None, t_rise=None):
        ud = np.ascontiguousarray(ubc_dof, dtype=np.int32)
        uv = np.ascontiguousarray(ubc_val, dtype=np.float64)
        fd = np.ascontiguousarray(fbc_dof, dtype=np.int32)
        fv = np.ascontiguousarray(fbc_val, dtype=np.float64)
        nrhs = 1 if amp is None else len(amp)
        a = None if amp is None else np.ascontiguousarray(amp, dtype=np.float64)
        t = None if t_rise is None else np.ascontiguousarray(t_rise, dtype=np.float64)
        if a is not None and t is None:
            t = np.zeros(nrhs)
        self._keep = (ud, uv, fd, fv, a, t)
        self.nrhs = nrhs
        self._check(lib().tdbem_march_init(
            self._h, a1, a2, len(ud), _i(ud), _d(uv), len(fd), _i(fd), _d(fv), nrhs,
            None if a is None else _d(a), None if t is None else _d(t)))

    def march_run(self, nsteps=None):
        n = self.numstep if nsteps is None else nsteps
        self._check(lib().tdbem_march_run(self._h, n))

    def march_enqueue(self, nsteps):
        """Queue nsteps on this context's stream without waiting."""
        self._check(lib().tdbem_march_enqueue(self._h, nsteps))

    def done(self):
        rc = lib().tdbem_query(self._h)
        if rc < 0:
            self._check(rc)
        return rc == 1

    @property
    def step_index(self):
        return lib().tdbem_march_step_index(self._h)

    def history(self, rhs=0):
        U = np.zeros((self.numstep, self.N))
        T = np.zeros((self.numstep, self.N))
        self._check(lib().tdbem_get_history(self._h, rhs, _d(U), _d(T)))
        return U, T

    def synchronize(self):
        self._check(lib().tdbem_synchronize(self._h))

    def set_precision(self, precision):
        """PRECISION_FP64 or PRECISION_MIXED (split-BF16 MFMA convolution) for march_init."""
        self._check(lib().tdbem_set_precision(self._h, int(precision)))

    def state(self, s, rhs=0):
        u = np.zeros(self.N)
        t = np.zeros(self.N)
        self._check(lib().tdbem_get_state(self._h, rhs, s, _d(u), _d(t)))
        return u, t

    TIMING_KEYS = ("store_ms", "init_ms", "march_ms", "lag_kernel_ms", "gather_ms", "conv_ms",
                   "solve_ms", "lag_launches", "conv_launches", "active_points", "store_nonzero",
                   "store_padded", "tile_subtiles", "reduce_ms", "tile_blocks", "tile_entries",
                   "active_levels", "pair_lags_near", "pair_lags_even", "pair_lags_odd",
                   "subtiles_near", "subtiles_even", "subtiles_odd", "entries_parity0",
                   "entries_parity1")

    def set_profiling(self, on=True, conv_only=False):
        """Event timing of the kernels (tdbem_set_profiling); conv_only times the history
        convolution alone (two events per step, no store-build timers or point counts)."""
        level = 0 if not on else (PROFILE_CONV if conv_only else PROFILE_ALL)
        self._check(lib().tdbem_set_profiling(self._h, level))

    def reset_timings(self):
        self._check(lib().tdbem_reset_timings(self._h))

    def timings(self):
        out = np.zeros(len(self.TIMING_KEYS))
        self._check(lib().tdbem_get_timings(self._h, _d(out)))
        return dict(zip(self.TIMING_KEYS, out.tolist()))


def solve_problem(p, amp=None, t_rise=None):
    """Full reference pipeline (main.cpp:89-128) on the GPU: lag store, march, histories."""
    ctx = Context.from_problem(p)
    ctx.march_init(p.a1, p.a2, p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val,
                   amp=[p.amp] if amp is None else amp,
                   t_rise=[p.t_rise] if t_rise is None else t_rise)
    ctx.march_run()
    out = [ctx.history(r) for r in range(ctx.nrhs)]
    ctx.close()
    return out


def mfma_selftest(A, B):
    """D = A B through one v_mfma_f32_16x16x32_bf16 with the lane layouts the mixed convolution
    uses (A 16x32, B 32x16, rounded to bf16)."""
    A = np.ascontiguousarray(A, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    D = np.zeros((16, 16), dtype=np.float32)
    fp = C.POINTER(C.c_float)
    rc = lib().tdbem_mfma_selftest(A.ctypes.data_as(fp), B.ctypes.data_as(fp),
                                   D.ctypes.data_as(fp))
    if rc:
        raise TdbemError(rc, lib().tdbem_last_error(None).decode())
    return D


def stream_copy(nbytes, launches=3):
    """Plain device copy of nbytes (k_stream_copy), `launches` times; returns the mean launch time
    in ms.  Its byte counters calibrate those of the other kernels."""
    ms = C.c_double(0.0)
    rc = lib().tdbem_stream_copy(C.c_size_t(nbytes), C.c_int(launches), C.byref(ms))
    if rc:
        raise TdbemError(rc, lib().tdbem_last_error(None).decode())
    return ms.value


def stream_read(nbytes, launches=3):
    """Read-only non-temporal stream over nbytes (k_stream_read); returns the mean launch time in
    ms: nbytes / time is the practical read bandwidth of this GPU."""
    ms = C.c_double(0.0)
    rc = lib().tdbem_stream_read(C.c_size_t(nbytes), C.c_int(launches), C.byref(ms))
    if rc:
        raise TdbemError(rc, lib().tdbem_last_error(None).decode())
    return ms.value


def device_memory():
    """(free, total) device bytes after a device synchronisation with the memory pool trimmed."""
    f, t = C.c_int64(0), C.c_int64(0)
    rc = lib().tdbem_device_memory(C.byref(f), C.byref(t))
    if rc:
        raise TdbemError(rc, lib().tdbem_last_error(None).decode())
    return f.value, t.value


def mfma_f64_selftest(A, B):
    """D = A B through one v_mfma_f64_16x16x4 with the lane layouts of the step solve
    (A 16x4, B 4x16)."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    D = np.zeros((16, 16))
    rc = lib().tdbem_mfma_f64_selftest(_d(A), _d(B), _d(D))
    if rc:
        raise TdbemError(rc, lib().tdbem_last_error(None).decode())
    return D


def dense_solve(A, B, inverse=False, pivoted=False):
    """X = A^{-1} B on the GPU with the step-0 solver: the march's LU and triangular solves
    (inverse=False) or the explicit inverse the march applies (inverse=True); pivoted=True forces
    partial pivoting.  B is n or n x nrhs."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    X = np.array(B, dtype=np.float64, order="C", copy=True)
    nrhs = 1 if X.ndim == 1 else X.shape[1]
    mode = (1 if inverse else 0) | (2 if pivoted else 0)
    rc = lib().tdbem_dense_solve(C.c_int(n), _d(A), C.c_int(nrhs), _d(X), C.c_int(mode))
    if rc:
        raise TdbemError(rc, lib().tdbem_last_error(None).decode())
    return X


def prepare_contexts(problems, loads, precision=PRECISION_MIXED, workers=8):
    """One context per problem with its set-up done (lag store, step-0 factorisations and
    inverse, march_init with the problem's boundary data), `workers` set-ups at a time: each
    context has its own HIP stream and the C ABI calls release the GIL, so the set-ups of many
    scenarios overlap on the device and on the host.  loads: per problem (amp, t_rise) arrays of
    the right-hand sides.  Returns the contexts in problem order; on an error the contexts
    already made are closed and the first error is raised."""
    from concurrent.futures import ThreadPoolExecutor

    def one(i):
        p = problems[i]
        amp, rise = loads[i]
        ctx = Context.from_problem(p)
        try:
            ctx.set_precision(precision)
            ctx.build_lag_store()
            ctx.march_init(p.a1, p.a2, p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val, amp=amp,
                           t_rise=rise)
        except BaseException:
            ctx.close()
            raise
        return ctx

    with ThreadPoolExecutor(max_workers=max(1, int(workers))) as ex:
        futs = [ex.submit(one, i) for i in range(len(problems))]
    ctxs, err = [], None
    for f in futs:
        try:
            ctxs.append(f.result())
        except BaseException as e:  # noqa: BLE001 - re-raised below after cleanup
            if err is None:
                err = e
    if err is not None:
        for c in ctxs:
            c.close()
        raise err
    return ctxs


class Batch:
    """Scenarios marched together (tdbem_batch_*): each context after a mixed-precision
    march_init with the same number of right-hand sides; results stay in the contexts.  Members
    may join (add) and leave (remove) between runs; a member cannot be closed or re-initialised
    until it has left."""

    def __init__(self, contexts=()):
        self.contexts = list(contexts)
        arr = (C.c_void_p * max(1, len(self.contexts)))(*[c._h for c in self.contexts])
        h = C.c_void_p()
        rc = lib().tdbem_batch_create(arr, len(self.contexts), C.byref(h))
        if rc:
            raise TdbemError(rc, lib().tdbem_last_error(None).decode())
        self._h = h

    def _check(self, rc):
        if rc:
            raise TdbemError(rc, lib().tdbem_last_error(None).decode())

    def add(self, ctx):
        """Admits a context (its first step is solved on its own stream if not yet done)."""
        self._check(lib().tdbem_batch_add(self._h, ctx._h))
        self.contexts.append(ctx)

    def remove(self, ctx):
        self._check(lib().tdbem_batch_remove(self._h, ctx._h))
        self.contexts.remove(ctx)

    def __len__(self):
        return lib().tdbem_batch_size(self._h)

    def tune(self):
        """Times the tunables on this batch and keeps the fastest."""
        rep = np.zeros(4)
        self._check(lib().tdbem_batch_tune(self._h, _d(rep)))
        return {"rounds": int(rep[0]), "solve_waves": int(rep[1]), "conv_ms": float(rep[2]),
                "solve_ms": float(rep[3])}

    def set_params(self, rounds, solve_waves):
        self._check(lib().tdbem_batch_set_params(self._h, int(rounds), int(solve_waves)))

    def run(self, steps_per_graph=16, max_replays=0):
        n = C.c_int(0)
        self._check(lib().tdbem_batch_run(self._h, int(steps_per_graph), int(max_replays),
                                          C.byref(n)))
        return n.value

    def close(self):
        if getattr(self, "_h", None):
            lib().tdbem_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Element:
    """TDElementT (src/TDElementT.h:16-115) as a reusable evaluator of (collocation point,
    element) pairs: kernel picks the element class (KERNEL_STEP: TDElementStepLinearT,
    KERNEL_HEAVISIDE: TDElementHeavisideLinearT, KERNEL_BSPLINE / KERNEL_RAMP:
    TDElementRampHeaviLinearT)."""

    def __init__(self, etype=1, kernel=KERNEL_STEP, nq=6, *, E, nu, rho):
        h = C.c_void_p()
        rc = lib().tdbem_element_create(int(etype), int(kernel), int(nq), float(E), float(nu),
                                        float(rho), C.byref(h))
        if rc:
            raise TdbemError(rc, lib().tdbem_element_last_error(None).decode())
        self._h = h
        self.ennd = 4 if etype == 1 else 8

    def _check(self, rc):
        if rc:
            raise TdbemError(rc, lib().tdbem_element_last_error(self._h).decode())

    def form(self, coords, collocation, singular, t, t_1):
        """SetCoords / SetCollocation / SetTime(t, t_1) / FormGH_Regular (singular -1) or
        FormGH_Singular(singular) / GetGHLinear for each pair.  coords [n][ennd][3] (or one
        [ennd][3]), collocation [n][3], singular [n].  Returns GE_1, GE_2, HE_1, HE_2
        [n][3][3 ennd] and C_1, C_2 [n][3][3] (without the leading n for a single pair)."""
        X = np.ascontiguousarray(coords, dtype=np.float64)
        single = X.ndim == 2
        X = X.reshape(-1, self.ennd, 3)
        n = X.shape[0]
        col = np.ascontiguousarray(collocation, dtype=np.float64).reshape(n, 3)
        sg = np.ascontiguousarray(np.broadcast_to(np.asarray(singular, dtype=np.int32), (n,)))
        out = [np.empty((n, 3, 3 * self.ennd)) for _ in range(4)] + \
              [np.empty((n, 3, 3)) for _ in range(2)]
        self._check(lib().tdbem_element_form(self._h, n, _d(X), _d(col), _i(sg), float(t),
                                             float(t_1), *[_d(o) for o in out]))
        return [o[0] for o in out] if single else out

    def form_device(self, coords, collocation, singular, t, t_1, out):
        """tdbem_element_form_device on torch tensors already on the device (float64 coords
        [n][ennd][3] and collocation [n][3], int32 singular [n], float64 out of
        n * (36 ennd + 18)); queued on the element's stream, see synchronize()."""
        n = collocation.shape[0]
        assert coords.numel() == n * 3 * self.ennd and out.numel() == n * (36 * self.ennd + 18)
        assert singular.numel() == n and singular.dtype.itemsize == 4
        self._check(lib().tdbem_element_form_device(
            self._h, n, C.c_void_p(coords.data_ptr()), C.c_void_p(collocation.data_ptr()),
            C.c_void_p(singular.data_ptr()), float(t), float(t_1), C.c_void_p(out.data_ptr())))

    def synchronize(self):
        self._check(lib().tdbem_element_synchronize(self._h))

    def close(self):
        if self._h:
            lib().tdbem_element_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class TwoRegionSolver:
    """SolverMRT (src/SolverMRT.cpp) over two FP64 region contexts (tdbem_mrt_*).

    iface1 / iface2: interface node ids (0-based) of each region; region 2's are matched to
    region 1's by coordinates.  max_step is SetMaxStep's value (the reference's default is 10).
    The region contexts must stay alive while the solver exists."""

    def __init__(self, region1, region2, iface1, iface2, max_step=10):
        self.regions = (region1, region2)
        i1 = np.ascontiguousarray(iface1, dtype=np.int32)
        i2 = np.ascontiguousarray(iface2, dtype=np.int32)
        if len(i1) != len(i2):
            raise ValueError("Interface mismatch, interface nodes must match one-to-one")
        h = C.c_void_p()
        rc = lib().tdbem_mrt_create(region1._h, region2._h, int(max_step), len(i1), _i(i1),
                                    _i(i2), C.byref(h))
        if rc:
            raise TdbemError(rc, lib().tdbem_last_error(None).decode())
        self._h = h
        self.n = lib().tdbem_mrt_num_unknowns(self._h)
        self.numstep = region1.numstep

    def _check(self, rc):
        if rc:
            raise TdbemError(rc, lib().tdbem_last_error(None).decode())

    def init(self, scheme=MRT_DIRECT, load_kind=MRT_LOAD_PULSE, amp=1.0, t_rise=0.0,
             fbc1=((), ()), fbc2=((), ())):
        """fbc1 / fbc2: (dofs, values) traction conditions of each region."""
        arrs = []
        for dof, val in (fbc1, fbc2):
            d = np.ascontiguousarray(dof, dtype=np.int32)
            v = np.ascontiguousarray(val, dtype=np.float64)
            if len(d) != len(v):
                raise ValueError("FBC dofs and values differ in length")
            arrs += [d, v]
        self._fbc = arrs
        self._check(lib().tdbem_mrt_init(self._h, int(scheme), int(load_kind), float(amp),
                                         float(t_rise), len(arrs[0]), _i(arrs[0]), _d(arrs[1]),
                                         len(arrs[2]), _i(arrs[2]), _d(arrs[3])))
        self.scheme = scheme

    def run(self, nsteps=None):
        self._check(lib().tdbem_mrt_run(self._h, int(nsteps if nsteps is not None else
                                                     self.numstep)))
        return lib().tdbem_mrt_step_index(self._h)

    def state(self, s):
        """(x, u, t) of step s: global unknowns, region displacements and tractions."""
        x, u, t = np.zeros(self.n), np.zeros(self.n), np.zeros(self.n)
        self._check(lib().tdbem_mrt_get_state(self._h, int(s), _d(x), _d(u), _d(t)))
        return x, u, t

    def history(self):
        """X, U, T [numstep + 1][n]."""
        X = np.zeros((self.numstep + 1, self.n))
        U, T = np.zeros_like(X), np.zeros_like(X)
        for s in range(self.numstep + 1):
            X[s], U[s], T[s] = self.state(s)
        return X, U, T

    def timings(self):
        a, b = C.c_double(), C.c_double()
        self._check(lib().tdbem_mrt_get_timings(self._h, C.byref(a), C.byref(b)))
        return {"init_ms": a.value, "run_ms": b.value}

    def close(self):
        if getattr(self, "_h", None):
            lib().tdbem_mrt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
