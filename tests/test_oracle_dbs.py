"""Pins the oracle's local 3x3 solvers (or_dbs.c, restating dbs.c:35-326):
converged (R, U) satisfy the Signorini-Coulomb law C(U, R) = 0 and U = B + W R;
the four solvers agree; separation / sticking / sliding have closed forms on
decoupled blocks; the 3x3 LU solve matches numpy and pivots."""
import ctypes as C

import numpy as np
import pytest

import oracle
from test_oracle_scf import residual

L = oracle.lib()
SSN, PG, DSF, PN = 2, 0, 1, 3


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def spd(rng, scale=1.0):
    A = rng.normal(size=(3, 3))
    return scale * (A @ A.T + 0.5 * np.eye(3))


def colmajor(M):
    return np.ascontiguousarray(M.T).ravel()


def test_solve3_matches_numpy_and_pivots():
    rng = np.random.default_rng(0)
    for _ in range(300):
        A = rng.normal(size=(3, 3))
        b = rng.normal(size=3)
        a, x = colmajor(A), b.copy()
        assert L.OR_Solve3(_p(a), _p(x)) == 0
        assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)
    # zero leading entry: only row pivoting makes LU possible
    A = np.array([[0.0, 2.0, 1.0], [3.0, 1.0, 0.0], [1.0, 0.0, 4.0]])
    b = np.array([1.0, 2.0, 3.0])
    a, x = colmajor(A), b.copy()
    assert L.OR_Solve3(_p(a), _p(x)) == 0
    assert np.allclose(x, np.linalg.solve(A, b))
    # exactly singular: reported (dgesv info > 0)
    a, x = colmajor(np.ones((3, 3))), b.copy()
    assert L.OR_Solve3(_p(a), _p(x)) != 0


def solve(solver, W, B, mu, R0=None, U0=None, eps=1e-12, maxiter=200, **kw):
    R0 = np.zeros(3) if R0 is None else R0
    U0 = np.zeros(3) if U0 is None else U0
    return oracle.diag_solve(solver, eps, maxiter, colmajor(W), B, np.zeros(3), U0, R0, mu, **kw)


@pytest.mark.parametrize("solver", [SSN, PG, DSF, PN])
def test_converged_solution_satisfies_contact_law(solver):
    rng = np.random.default_rng(10 + solver)
    ok = 0
    for _ in range(200):
        W = spd(rng, 1e-3)
        B = rng.normal(size=3)
        mu = rng.uniform(0.1, 1.0)
        it, R, U = solve(solver, W, B, mu, maxiter=5000 if solver in (PG, DSF) else 200, eps=1e-12)
        if it < 0 or not np.all(np.isfinite(R)):
            continue
        ok += 1
        assert np.allclose(U, B + W @ R, atol=1e-9 * max(1.0, np.abs(B).max()))
        Cv = residual(U, R, mu)
        assert np.abs(Cv).max() <= 1e-6 * max(1.0, np.abs(R).max(), np.abs(U).max()), (solver, Cv)
    assert ok >= 150


def test_four_solvers_agree():
    rng = np.random.default_rng(20)
    for _ in range(100):
        W = spd(rng, 1e-2)
        B = rng.normal(size=3)
        mu = rng.uniform(0.1, 1.0)
        sols = []
        for s in (SSN, PG, DSF, PN):
            it, R, U = solve(s, W, B, mu, eps=1e-13, maxiter=20000)
            if it >= 0 and np.all(np.isfinite(R)):
                sols.append(R)
        assert len(sols) >= 3
        for R in sols[1:]:
            assert np.allclose(R, sols[0], rtol=1e-6, atol=1e-6 * max(1.0, np.abs(sols[0]).max()))


def test_separation_closed_form():
    W = np.diag([2.0, 2.0, 1.0])
    B = np.array([0.3, -0.1, 0.5])          # normal velocity opening
    it, R, U = solve(SSN, W, B, 0.4)
    assert np.allclose(R, 0.0, atol=1e-14) and np.allclose(U, B)


def test_positive_gap_early_exit():
    W = np.diag([2.0, 2.0, 1.0])
    B = np.array([0.3, -0.1, -0.5])
    for s in (SSN, PG, DSF):
        it, R, U = solve(s, W, B, 0.4, R0=np.ones(3), gap=1e-3)
        assert it == 0 and np.all(R == 0.0) and np.all(U == B)


def test_sticking_closed_form():
    w, wn = 3.0, 1.5
    W = np.diag([w, w, wn])
    B = np.array([0.0, 0.0, -0.7])
    it, R, U = solve(SSN, W, B, 0.5)
    assert np.allclose(R, [0.0, 0.0, 0.7 / wn], rtol=1e-12, atol=1e-14)
    assert np.allclose(U, 0.0, atol=1e-12)


def test_sliding_closed_form():
    w, wn, mu = 2.0, 1.0, 0.3
    W = np.diag([w, w, wn])
    B = np.array([1.0, 0.5, -0.4])           # |B_T| > mu w R_N: slides
    rn = 0.4 / wn
    bt = B[:2]
    rt = -mu * rn * bt / np.linalg.norm(bt)
    it, R, U = solve(SSN, W, B, mu)
    assert np.allclose(R, [rt[0], rt[1], rn], rtol=1e-10)
    assert np.allclose(U[:2] / np.linalg.norm(U[:2]), bt / np.linalg.norm(bt))
    assert abs(U[2]) < 1e-12


def test_newton_iteration_counts_are_small():
    rng = np.random.default_rng(30)
    its = []
    for _ in range(200):
        W = spd(rng, 1e-3)
        B = rng.normal(size=3)
        it, R, U = solve(SSN, W, B, rng.uniform(0.1, 1.0), eps=1e-10)
        its.append(it)
    assert np.median(its) <= 10 and max(its) < 200
