"""Pins the oracle's NEWTON_SOLVER (or_nts.c, restating nts.c:937-1379 with the
flexible GMRES of ext/krylov/flexgmres.c): the Krylov routine against
numpy.linalg.solve on dense systems; the Newton solve by the contact law at its
converged state, by decoupled rows (R_N = m g), by agreement with the
Gauss-Seidel oracle's invariants, by its iteration and product bounds and by
the Gauss-Seidel fallback."""
import numpy as np
import pytest

import oracle
from scenes import lattice
from test_oracle_scf import residual

H = 1e-3


def rand_system(n, seed, dominance=0.0):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(n, n)) + dominance * np.eye(n)
    return A, rng.normal(size=n)


@pytest.mark.parametrize("n,seed", [(5, 1), (8, 2), (20, 3)])
def test_gmres_is_exact_after_n_iterations(n, seed):
    """Without a tolerance GMRES runs every iteration; after n of them (n <= k_dim
    = 20) the Krylov space is the whole space and x solves A x = b."""
    A, b = rand_system(n, seed)
    it, x, rn, mv = oracle.flexgmres_dense(A, b, maxiter=n, atol=0.0)
    xs = np.linalg.solve(A, b)
    assert it == n and mv == n + 1          # the initial residual and one product per iteration
    np.testing.assert_allclose(x, xs, rtol=0, atol=1e-9 * np.abs(xs).max())
    assert rn <= 1e-9 * np.linalg.norm(b)


def test_gmres_preconditioned_restarted_and_trivial():
    A, b = rand_system(30, 4, dominance=12.0)
    xs = np.linalg.solve(A, b)
    tol = 1e-12 * np.linalg.norm(b)
    # right preconditioning with the inverse diagonal; restarted every 3 iterations
    it, x, rn, mv = oracle.flexgmres_dense(A, b, P=np.diag(1.0 / np.diag(A)), kdim=3, maxiter=200, atol=tol)
    assert it < 200 and np.linalg.norm(b - A @ x) <= tol
    np.testing.assert_allclose(x, xs, rtol=0, atol=1e-10 * np.abs(xs).max())
    # the exact inverse as the preconditioner: one iteration
    it, x, rn, mv = oracle.flexgmres_dense(A, b, P=np.linalg.inv(A), maxiter=10, atol=1e-10 * np.linalg.norm(b))
    assert it == 1
    np.testing.assert_allclose(x, xs, rtol=0, atol=1e-9 * np.abs(xs).max())
    # from the solution: the residual is zero before the first iteration
    it, x, rn, mv = oracle.flexgmres_dense(A, A @ np.ones(30), x0=np.ones(30), maxiter=10, atol=0.0)
    assert it in (0, 10) and np.abs(x - 1.0).max() < 1e-12
    it, x, rn, mv = oracle.flexgmres_dense(A, np.zeros(30), maxiter=10)
    assert it == 0 and np.all(x == 0.0)


@pytest.fixture(scope="module")
def system():
    s = lattice(216, seed=3)
    b, c = s.bodies(), s.contacts()
    O = oracle.assemble(oracle.bodies_from(b), oracle.contacts_from(c), H)
    return O, oracle.System(O), b, c


# (linver, meritval, bound on the local residual and on the body velocities
# against the Gauss-Seidel solution, relative): DIAG converges linearly
@pytest.mark.parametrize("linver,meritval,ctol,vtol", [("GMRES", 1e-12, 1e-8, 1e-2), ("DIAG", 1e-8, 1e-5, 0.1)])
def test_newton_converges_to_the_contact_law(system, linver, meritval, ctol, vtol):
    O, S = system[:2]
    n = O["n"]
    rep, R, U, mer, mv = oracle.newton_solve(S, oracle.nt_param(linver=linver, maxiter=1000, meritval=meritval),
                                             np.zeros(3 * n))
    assert rep.converged and not rep.gs_fallback and rep.merit <= meritval
    assert len(mer) == rep.iters and mer[-1] == rep.merit and np.all(np.diff(mv) >= 0)
    # U = W R + B is what the solve leaves; the local residual C(U, R) is small
    m, Um, _ = oracle.merit(S, R.ravel(), U.ravel(), 1)
    np.testing.assert_allclose(Um, U, rtol=0, atol=1e-12 * np.abs(U).max())
    assert m == pytest.approx(rep.merit, rel=1e-12)
    Cv = np.array([residual(U[i], R[i], O["mu"][i]) for i in range(n)])
    assert np.abs(Cv).max() <= ctol * max(np.abs(R).max(), np.abs(U).max())
    # the Gauss-Seidel oracle to a small merit: frictional packings need not have
    # unique reactions, but the body velocities agree as far as the merits allow
    b, c = system[2:]
    nc, col = oracle.colour(O["rowptr"], O["col"])
    rc, grep, Rg, Ug, _, _ = oracle.gs_solve(S, oracle.gs_param(1e-10, 20000, meritval=1e-10),
                                             oracle.colour_order(col, nc)[0], np.zeros(3 * n), np.zeros(3 * n))
    vn, _ = oracle.body_update(oracle.bodies_from(b), oracle.contacts_from(c), R.ravel(), H)
    vg, _ = oracle.body_update(oracle.bodies_from(b), oracle.contacts_from(c), Rg.ravel(), H)
    assert np.abs(vn - vg).max() <= vtol * np.abs(vg).max()


def decoupled(n, m=1.0, g=9.81, mu=0.5):
    """n spheres each resting on its own plane: W_ii = (h/m) diag(7/2, 7/2, 1), B = (0, 0, -h g)."""
    Wd = np.tile((H / m * np.diag([3.5, 3.5, 1.0])).T.ravel(), n)
    A = np.tile((m / H * np.diag([1 / 3.5, 1 / 3.5, 1.0])).T.ravel(), n)
    B = np.tile([0.0, 0.0, -H * g], n)
    d = dict(n=n, rowptr=np.zeros(n + 1, np.int32), col=np.zeros(0, np.int32), W=np.zeros(0), Wd=Wd, A=A, B=B,
             V=np.zeros(3 * n), rho=np.full(n, m / H), mu=np.full(n, mu), rest=np.zeros(n), coh=np.zeros(n),
             gap=np.zeros(n), free_energy=0.5 * n * m * H * g * g, step=H)
    return oracle.System(d)


@pytest.mark.parametrize("linver", ["GMRES", "DIAG"])
def test_decoupled_rows_carry_the_weight(linver):
    S = decoupled(4)
    rep, R, U, mer, mv = oracle.newton_solve(S, oracle.nt_param(linver=linver, theta=1.0), np.zeros(12))
    assert rep.converged
    np.testing.assert_allclose(R, np.tile([0.0, 0.0, 9.81], (4, 1)), rtol=0, atol=1e-9)
    np.testing.assert_allclose(U, 0.0, atol=1e-12)


def test_bounds_and_fallback(system):
    O, S = system[:2]
    n = O["n"]
    # the product bound ends the solve after the first GMRES step
    rep, R, U, mer, mv = oracle.newton_solve(S, oracle.nt_param(maxmatvec=2, gsflag="OFF"), np.zeros(3 * n))
    assert rep.iters == 1 and mv[0] == rep.matvecs and 2 <= rep.matvecs <= 12 and not rep.converged
    # GMRES stops at epsilon |b| (0.25 by default) or after linmaxiter iterations;
    # unconverged, it takes the initial residual and one product per iteration
    # (the closing true residual only follows a step below the tolerance)
    rep, R, U, mer, mv = oracle.newton_solve(S, oracle.nt_param(maxiter=5, epsilon=1e-30, linmaxiter=7,
                                                                gsflag="OFF"), np.zeros(3 * n))
    assert list(mv) == [8 * (k + 1) for k in range(5)]
    # unconverged with gsflag: R back to R0, then GAUSS_SEIDEL_SOLVER(1.0, maxiter, meritval, ...)
    R0 = np.random.default_rng(5).normal(scale=1e-3, size=3 * n)
    p = oracle.nt_param(maxiter=3, meritval=1e-9)
    rep, R, U, mer, mv = oracle.newton_solve(S, p, R0)
    assert rep.iters == 3 and rep.gs_fallback and not rep.converged
    # the same Gauss-Seidel solve run directly from R0 and the Newton's last U
    rep2, _, Ulast, _, _ = oracle.newton_solve(S, oracle.nt_param(maxiter=3, meritval=1e-9, gsflag="OFF"), R0)
    gp = oracle.gs_param(1.0, 3, meritval=1e-9, diagepsilon=1e-9, diagmaxiter=100)
    rc, grep, Rg, Ug, _, _ = oracle.gs_solve(S, gp, np.arange(n, dtype=np.int32), R0, Ulast.ravel())
    np.testing.assert_array_equal(R, Rg)
    assert rep.gs_iters == grep.iters
