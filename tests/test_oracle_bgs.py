"""Pins the oracle's Gauss-Seidel driver and merit (or_bgs.c, restating
bgs.c:913-1069 and mrf.c:35-184): at convergence every row satisfies its local
contact law; the merit of an exact solution is 0 and of R = 0 has a closed form;
decoupled rows converge in one sweep; list and colour orders reach the same
invariants (merit, per-body impulse) while their reactions may differ by the
null space of a hyperstatic W; failure actions and error codes."""
import numpy as np
import pytest

import oracle
from scenes import lattice
from test_oracle_scf import residual, soc_textbook

H = 1e-3


@pytest.fixture(scope="module")
def system():
    s = lattice(216, seed=3)
    b, c = s.bodies(), s.contacts()
    O = oracle.assemble(oracle.bodies_from(b), oracle.contacts_from(c), H)
    return b, c, O, oracle.System(O)


def colour_order(O):
    nc, col = oracle.colour(O["rowptr"], O["col"])
    return oracle.colour_order(col, nc)[0]


def local_residuals(O, R, U):
    n = O["n"]
    out = []
    for i in range(n):
        out.append(residual(U[i], R[i], O["mu"][i]))
    return np.array(out)


def test_converged_rows_satisfy_the_contact_law(system):
    b, c, O, S = system
    n = O["n"]
    # a hyperstatic packing: the relative change stalls near 1e-6 (reactions drift in the
    # null space of W) while the merit keeps falling, so the test runs to 1e-6
    rc, rep, R, U, rer, mer = oracle.gs_solve(S, oracle.gs_param(1e-6, 20000), colour_order(O),
                                              np.zeros(3 * n), np.zeros(3 * n))
    # local failures are recovered by the fallback chain (error DIAGONAL_*, not DIVERGED)
    assert rc == 0 and rep.iters < 20000 and rep.error in (0, 2, 3)
    # U = B + W R up to the last sweep's change of the neighbours (the merit recomputes it)
    m, Um, _ = oracle.merit(S, R.ravel(), U.ravel(), 1)
    assert np.abs(Um - U).max() <= 1e-4 * np.abs(U).max()
    Cv = local_residuals(O, R, Um)
    assert np.abs(Cv).max() <= 1e-6 * max(np.abs(R).max(), np.abs(Um).max())
    assert m < 1e-10


def test_merit_closed_form_at_zero_reaction(system):
    b, c, O, S = system
    n = O["n"]
    m, U, cm = oracle.merit(S, np.zeros(3 * n), np.zeros(3 * n), 1)
    num = den = 0.0
    for i in range(n):
        Bi = O["B"][3 * i:3 * i + 3]
        Ai = O["A"][9 * i:9 * i + 9].reshape(3, 3).T
        F = np.array([Bi[0], Bi[1], Bi[2] + O["mu"][i] * np.linalg.norm(Bi[:2])])
        Ci = -soc_textbook(O["mu"][i], -F)
        num += Ci @ Ai @ Ci
        den += Bi @ Ai @ Bi
    assert np.allclose(U.ravel(), O["B"])
    assert abs(m - num / den) <= 1e-12 * max(abs(m), 1e-300)
    assert abs(np.sum(cm) - 2 * m) <= 1e-10 * max(abs(m), 1e-300)


def test_decoupled_rows_converge_in_one_sweep():
    """Spheres resting on the floor, far apart: no W_ij, so the first sweep solves
    every row and the second sweep changes nothing (error 0)."""
    import solfec
    s = solfec.SOLFEC("DYNAMIC", H, "/tmp/solfec_test")
    solfec.SURFACE_MATERIAL(s, friction=0.3)
    solfec.GRAVITY(s, (0, 0, -9.81))
    bulk = solfec.BULK_MATERIAL(s, density=2000.0)
    for k in range(5):
        solfec.BODY(s, "RIGID", solfec.SPHERE((0.3 * k, 0.0, 0.0499), 0.05, 1, 1), bulk)
    solfec.BODY(s, "OBSTACLE", solfec.PLANE((0, 0, 0), (0, 0, 1), 0), bulk)
    s.step_begin()
    b, c = s.bodies(), s.contacts()
    O = oracle.assemble(oracle.bodies_from(b), oracle.contacts_from(c), H)
    assert O["n"] == 5 and O["nnz"] == 0
    rc, rep, R, U, rer, mer = oracle.gs_solve(oracle.System(O), oracle.gs_param(1e-8, 100), oracle.list_order(5),
                                              np.zeros(15), np.zeros(15))
    assert rep.iters == 2 and rer[1] == 0.0
    # resting: R is the mean force over the step (W carries h), so R_N = m g cancels gravity
    m = 2000.0 * 4.0 / 3.0 * np.pi * 0.05 ** 3
    assert np.allclose(R[:, 2], m * 9.81, rtol=1e-10)
    assert np.allclose(R[:, :2], 0.0, atol=1e-15)


def test_list_and_colour_orders_agree_on_invariants(system):
    b, c, O, S = system
    n = O["n"]
    par = oracle.gs_param(1e-6, 20000)
    out = {}
    for name, order in (("list", oracle.list_order(n)), ("colour", colour_order(O))):
        rc, rep, R, U, rer, mer = oracle.gs_solve(S, par, order, np.zeros(3 * n), np.zeros(3 * n))
        velo, force = oracle.body_update(oracle.bodies_from(b), oracle.contacts_from(c), R.ravel(), H)
        out[name] = (R, velo, oracle.merit(S, R.ravel(), U.ravel(), 1)[0])
    (Rl, vl, ml), (Rc, vc, mc) = out["list"], out["colour"]
    assert ml < 1e-9 and mc < 1e-9
    # frictional packings need not have unique solutions; the two orders must still
    # give the same body velocities to within the accuracy the stopping rule allows
    assert np.abs(vl - vc).max() <= 1e-2 * np.abs(vl).max()


def test_failure_exit_and_error_codes(system):
    b, c, O, S = system
    n = O["n"]
    order = colour_order(O)
    # maxiter = 1 with a tiny epsilon: DIVERGED, EXIT throws
    rc, rep, *_ = oracle.gs_solve(S, oracle.gs_param(1e-12, 1, failure=1), order, np.zeros(3 * n), np.zeros(3 * n))
    assert rc == 1 and rep.error == 1
    rc, rep, *_ = oracle.gs_solve(S, oracle.gs_param(1e-12, 1, failure=0), order, np.zeros(3 * n), np.zeros(3 * n))
    assert rc == 0 and rep.error == 1
    # diagmaxiter = 1 with a tight diagepsilon: local divergence, fallbacks under CONTINUE
    rc, rep, *_ = oracle.gs_solve(S, oracle.gs_param(1e-4, 50, diagepsilon=1e-300, diagmaxiter=1), order,
                                  np.zeros(3 * n), np.zeros(3 * n))
    assert rc == 0 and rep.diag_fail > 0 and rep.error in (1, 2, 3)
    rc, rep, *_ = oracle.gs_solve(S, oracle.gs_param(1e-4, 50, failure=1, diagepsilon=1e-300, diagmaxiter=1), order,
                                  np.zeros(3 * n), np.zeros(3 * n))
    assert rc == 1 and rep.error == 2


def test_reverse_alternation_converges(system):
    b, c, O, S = system
    n = O["n"]
    rc, rep, R, U, rer, mer = oracle.gs_solve(S, oracle.gs_param(1e-6, 20000, reverse=1), colour_order(O),
                                              np.zeros(3 * n), np.zeros(3 * n))
    assert rc == 0 and rep.iters < 20000
    assert oracle.merit(S, R.ravel(), U.ravel(), 1)[0] < 1e-9
