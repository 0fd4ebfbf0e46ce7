"""The step-0 dense solver (the KSPSolve stand-in, src/SolverStepLinearT.cpp:503-504) on its own,
through the C ABI, against numpy's LAPACK solve on the same seeded matrices: the blocked LU
without interchanges (taken when every multiplier has |l| <= 1, as on the march's step-0
matrices), the fallback to partial pivoting (general matrices), forced pivoting, and the explicit
inverse the march applies."""
import numpy as np
import pytest

from tdbem import api

pytestmark = pytest.mark.gpu


def _case(n, seed, kind):
    rng = np.random.default_rng(seed)
    if kind == "dominant":
        # like the step-0 BEM matrices: partial pivoting keeps every row in place
        A = rng.standard_normal((n, n)) / np.sqrt(n) * 0.3 + np.diag(rng.uniform(0.5, 1.5, n))
    else:
        A = rng.standard_normal((n, n))
    if kind == "permuted":
        # a tiny diagonal: partial pivoting has to move nearly every row
        A = A[rng.permutation(n)] * 1e-3 + np.eye(n)[rng.permutation(n)] * 3.0
        A[np.arange(n), np.arange(n)] = 1e-12
    B = rng.standard_normal((n, 3))
    return A, B


def _check(A, B, X, tol_res=1e-14):
    ref = np.linalg.solve(A, B)
    res = np.linalg.norm(A @ X - B) / (np.linalg.norm(A) * np.linalg.norm(X))
    assert res <= tol_res, res
    # both are backward stable: forward error within the condition number of rounding
    cond = np.linalg.cond(A)
    err = np.linalg.norm(X - ref) / np.linalg.norm(ref)
    assert err <= 1e-15 * cond * A.shape[0] ** 0.5 + 1e-13, (err, cond)


@pytest.mark.parametrize("n,kind", [(7, "dominant"), (64, "dominant"), (1100, "dominant"),
                                    (5000, "dominant"), (7, "normal"), (300, "permuted"),
                                    (1100, "normal"), (6000, "permuted")])
def test_dense_solve_matches_lapack(gpu, n, kind):
    A, B = _case(n, n, kind)
    print(f"dense solve n={n} {kind}", flush=True)
    _check(A, B, api.dense_solve(A, B))


@pytest.mark.parametrize("n", [65, 2500])
def test_forced_pivoting_matches_lapack(gpu, n):
    A, B = _case(n, 11 + n, "dominant")
    print(f"dense solve pivoted n={n}", flush=True)
    X = api.dense_solve(A, B, pivoted=True)
    _check(A, B, X)
    np.testing.assert_allclose(X, api.dense_solve(A, B), rtol=0, atol=1e-13 * np.abs(X).max())


@pytest.mark.parametrize("n,kind", [(100, "normal"), (2500, "dominant"), (5200, "normal")])
def test_dense_inverse_matches_lapack(gpu, n, kind):
    A, B = _case(n, 7 + n, kind)
    print(f"dense inverse n={n} {kind}", flush=True)
    X = api.dense_solve(A, B[:, :2], inverse=True)
    _check(A, B[:, :2], X, tol_res=1e-12)


@pytest.mark.parametrize("pivoted", [False, True])
def test_dense_solve_singular_is_an_error(gpu, pivoted):
    A = np.random.default_rng(3).standard_normal((40, 40))
    A[:, 17] = 0.0
    with pytest.raises(api.TdbemError) as e:
        api.dense_solve(A, np.ones(40), pivoted=pivoted)
    assert e.value.code == api.ERR_NUMERIC
