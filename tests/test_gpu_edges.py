"""GPU parity at the edges the other tests leave: the golden fixtures written from the
reference's own compiled classes (compared with the GPU directly, not through the oracle), the
10- and 20-point shape rules (src/ShapeFunctionT.cpp:71-169, with the reference's 20-point
abscissa slip at :121), the 8-node element (EType 2) through the lag store and the march
(src/TDElementStepLinearT.cpp:149-386 feeding UpdateGHLinear), and the errors of the C ABI
(argument checks, call order, batch membership).  Relative Frobenius error <= 1e-10 as in
tests/test_gpu_parity.py."""
from pathlib import Path

import numpy as np
import pytest

import helpers
import oracle_lib as O
from tdbem import api, configs

pytestmark = pytest.mark.gpu
TOL = 1e-10
GOLD = Path(__file__).resolve().parent / "golden"


def _announce(name):
    print(f"[gpu case] {name}", flush=True)


def _ctx(p, **kw):
    args = dict(E=p.E, nu=p.nu, rho=p.rho, dt=p.dt, numstep=p.numstep, nq=p.nq, etype=p.etype,
                if_ext=p.if_ext, kernel=p.kernel)
    args.update(kw)
    return api.Context(p.coords, p.ien, **args)


@pytest.mark.parametrize("name", ["plane1e", "bar_h05"])
def test_gpu_matches_reference_golden_assembly(gpu, name):
    """tests/golden/assemble_*.npz: G_H_Driver's four row blocks written by the reference's
    compiled TDElementStepLinearT + the assembly loop, at several steps."""
    g = np.load(GOLD / f"assemble_{name}.npz")
    upper = int(g["upper"])
    steps = [int(n) for n in g["steps"]]
    _announce(f"golden assembly {name} steps {steps}")
    with api.Context(g["coords"], g["ien"], E=float(g["E"]), nu=float(g["nu"]),
                     rho=float(g["rho"]), dt=float(g["dt"]), numstep=max(steps) + 1,
                     if_ext=int(g["if_ext"])) as ctx:
        got = [ctx.assemble_step(n, 0, upper) for n in steps]
    for m in range(4):
        err = helpers.rel_fro_set([x[m] for x in got], [r[m] for r in g["GH"]])
        assert err <= TOL, (name, m, err)


@pytest.mark.parametrize("nq", [10, 20])
def test_assemble_step_higher_rules(gpu, nq):
    p = helpers.small_bar(h=1.0, numstep=12)
    p.nq = nq
    _announce(f"assemble_step nq={nq}")
    with _ctx(p) as ctx:
        got = [ctx.assemble_step(n) for n in (1, 2, 5)]
    ref = [O.assemble_step(4, 1, nq, p.E, p.nu, p.rho, p.coords, p.ien, 0, 0, p.nnd - 1,
                           n * p.dt, p.dt, n) for n in (1, 2, 5)]
    for m in range(4):
        assert helpers.rel_fro_set([g[m] for g in got], [r[m] for r in ref]) <= TOL


@pytest.mark.parametrize("nq", [10, 20])
def test_lag_store_and_march_higher_rules(gpu, nq):
    p = helpers.small_bar(h=1.0, numstep=16)
    p.nq = nq
    _announce(f"lag store + march nq={nq}")
    with _ctx(p) as ctx:
        K = ctx.max_step
        ctx.march_init(p.a1, p.a2, p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val)
        ctx.march_run()
        U, T = ctx.history()
        Kc = min(K, p.numstep)
        lags = [ctx.get_lag(k) for k in range(Kc + 1)]
    MG, MH = O.build_lags(4, 1, nq, p.E, p.nu, p.rho, p.coords, p.ien, 0, p.dt, p.numstep, K)
    # the 20-point rule (with the reference's repeated abscissa) reaches closer to the singular
    # corner, where the oracle's static-tail noise eps t^2 / r^2 is largest: parity over the
    # entries the shell store holds, the noise outside bounded (tests/test_gpu_configs.py)
    for m, M in ((0, MG), (1, MH)):
        inside, outside = helpers.shell_parity([l[m] for l in lags], list(M[:Kc + 1]))
        print(f"nq={nq} lags {'GH'[m]}: in shell {inside:.2e}, oracle noise outside {outside:.2e}")
        assert inside <= TOL and outside <= 1e-9
    bc = (p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val)
    Ur, Tr = O.march(helpers.shell_masked(MG[:Kc + 1], [g for g, _ in lags]),
                     helpers.shell_masked(MH[:Kc + 1], [h for _, h in lags]), p.numstep, p.dt,
                     *bc)
    eu, et = helpers.rel_fro(U, Ur), helpers.rel_fro(T, Tr)
    print(f"nq={nq} march: u {eu:.2e} t {et:.2e}")
    # at nq = 20 the oracle and the reference's own compiled classes already differ by up to
    # 2e-10 in assembled rows (tests/test_oracle_ref.py); the march carries that level
    tol = TOL if nq < 20 else 3e-10
    assert eu <= tol and et <= tol


def test_quadratic_elements_lag_store_and_march(gpu):
    """EType 2 (8-node serendipity quads, singular pairs through bilinear sub-elements): the bar
    h = 1 as 8-node elements, clamped at z = 0 and pulled at z = 4."""
    b = configs.bar(h=1.0, numstep=14)
    p = helpers.small_quadratic_bar(numstep=14)
    # boundary data of the 8-node mesh: every node on z = 0 clamped, t_z = 1 on z = 4
    z = p.coords[:, 2]
    bot = np.where(np.abs(z) < 1e-12)[0]
    top = np.where(np.abs(z - 4.0) < 1e-12)[0]
    ubc = np.sort(np.concatenate([3 * bot + d for d in range(3)])).astype(np.int32)
    fbc = np.sort(3 * top + 2).astype(np.int32)
    assert p.nel == b.nel and p.ien.shape[1] == 8
    _announce(f"8-node bar: lag store + march nnd={p.nnd}")
    with _ctx(p) as ctx:
        K = ctx.max_step
        ctx.march_init(0.25, 0.25, ubc, np.zeros(len(ubc)), fbc, np.ones(len(fbc)))
        ctx.march_run()
        U, T = ctx.history()
        Kc = min(K, p.numstep)
        lags = [ctx.get_lag(k) for k in range(Kc + 1)]
    MG, MH = O.build_lags(4, 2, 6, p.E, p.nu, p.rho, p.coords, p.ien, 0, p.dt, p.numstep, K)
    assert helpers.rel_fro_set([g for g, _ in lags], list(MG[:Kc + 1])) <= TOL
    assert helpers.rel_fro_set([h for _, h in lags], list(MH[:Kc + 1])) <= TOL
    Ur, Tr = O.march(MG, MH, p.numstep, p.dt, ubc, np.zeros(len(ubc)), fbc, np.ones(len(fbc)))
    assert np.linalg.norm(Ur) > 0
    assert helpers.rel_fro(U, Ur) <= TOL and helpers.rel_fro(T, Tr) <= TOL


def _code(fn):
    with pytest.raises(api.TdbemError) as e:
        fn()
    return e.value.code


def test_c_abi_argument_and_order_errors(gpu):
    p = helpers.small_bar(h=1.0, numstep=8)
    _announce("C ABI errors")
    # create: unsupported rule, element type, kernel; bad material
    assert _code(lambda: _ctx(p, nq=7)) == api.ERR_ARG
    assert _code(lambda: _ctx(p, etype=3)) == api.ERR_ARG
    assert _code(lambda: _ctx(p, kernel=3)) == api.ERR_ARG
    assert _code(lambda: _ctx(p, nu=0.5)) == api.ERR_ARG
    assert _code(lambda: api.Context(p.coords, p.ien + p.nnd, E=1, nu=0.25, rho=1, dt=0.1,
                                     numstep=3)) == api.ERR_ARG
    with _ctx(p) as ctx:
        # call order
        assert _code(lambda: ctx.march_run()) == api.ERR_STATE
        assert _code(lambda: ctx.get_lag(0)) == api.ERR_STATE
        assert _code(lambda: ctx.history()) == api.ERR_STATE
        # alpha range (src/SolverStepLinearT.cpp:317-322) and right-hand-side counts
        bc = (p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val)
        for a1, a2 in ((-0.1, 0.2), (0.6, 0.5), (0.2, 1.0)):
            assert _code(lambda: ctx.march_init(a1, a2, *bc)) == api.ERR_ARG
        assert _code(lambda: ctx.march_init(0.25, 0.25, *bc, amp=[1, 1, 1])) == api.ERR_ARG
        ctx.set_precision(api.PRECISION_MIXED)
        assert _code(lambda: ctx.march_init(0.25, 0.25, *bc, amp=[1.0] * 8)) == api.ERR_ARG
        ctx.set_precision(api.PRECISION_FP64)
        assert _code(lambda: ctx.set_precision(7)) == api.ERR_ARG
        assert _code(lambda: ctx.march_init(0.25, 0.25, [10 ** 6], [0.0], [], [])) == api.ERR_ARG
        # lag and row ranges
        ctx.build_lag_store()
        assert _code(lambda: ctx.get_lag(ctx.max_step + 1)) == api.ERR_ARG
        assert _code(lambda: ctx.get_lag_rows(0, 3, 2)) == api.ERR_ARG
        assert _code(lambda: ctx.get_lag_rows(0, 0, p.nnd)) == api.ERR_ARG
        assert _code(lambda: ctx.assemble_step(0)) == api.ERR_ARG
        assert _code(lambda: ctx.state(p.numstep + 3)) == api.ERR_STATE
    # the assembly-only kernels have no march
    with _ctx(p, kernel=api.KERNEL_HEAVISIDE) as ctx:
        assert _code(lambda: ctx.build_lag_store()) == api.ERR_ARG


def test_batch_members_are_protected(gpu):
    """A context in a batch cannot be freed or re-initialised (the batch's graph reads its march
    buffers and tile store): TDBEM_ERR_STATE from the host checks, until it has left."""
    p = helpers.small_bar(h=1.0, numstep=10)
    amp, rise = configs.rhs_set(16, p.dt, seed=1)
    _announce("batch membership guards")
    ctx = api.Context.from_problem(p)
    ctx.set_precision(api.PRECISION_MIXED)
    bc = (p.ubc_dof, p.ubc_val, p.fbc_dof, p.fbc_val)
    ctx.march_init(p.a1, p.a2, *bc, amp=amp, t_rise=rise)
    b = api.Batch([ctx])
    assert len(b) == 1
    assert _code(lambda: ctx.close()) == api.ERR_STATE
    assert _code(lambda: ctx.march_init(p.a1, p.a2, *bc, amp=amp, t_rise=rise)) == api.ERR_STATE
    assert _code(lambda: ctx.march_run()) == api.ERR_STATE
    assert _code(lambda: ctx.build_lag_store()) == api.ERR_STATE
    assert _code(lambda: ctx.set_precision(api.PRECISION_FP64)) == api.ERR_STATE
    assert _code(lambda: ctx.set_max_step(3)) == api.ERR_STATE
    assert _code(lambda: b.add(ctx)) == api.ERR_ARG
    b.run(steps_per_graph=4)
    assert ctx.step_index == p.numstep - 1
    b.remove(ctx)
    assert len(b) == 0 and b.run() == 0
    ctx.march_init(p.a1, p.a2, *bc, amp=amp, t_rise=rise)   # allowed again
    b.close()
    ctx.close()
