"""Contact detection and the rigid time step: the oracle restatement (or_dom.c)
against closed forms, and the host layer's grid detection and integration
(the product's CPU loop, which the device loop reproduces) against the oracle's
brute force.  CPU only."""
import numpy as np
import pytest

import oracle
import solfec
from scenes import lattice


def rodrigues(O):
    """exp([O]x) from numpy's sin and cos, column-major 9-vector."""
    t = np.linalg.norm(O)
    K = np.array([[0, -O[2], O[1]], [O[2], 0, -O[0]], [-O[1], O[0], 0]])
    if t == 0:
        return np.eye(3).T.ravel()
    M = np.eye(3) + np.sin(t) / t * K + (1 - np.cos(t)) / t ** 2 * (K @ K)
    return M.T.ravel()


# ---------------- oracle against closed forms ----------------

def test_sphere_pair_closed_forms():
    ca, cb = np.array([0.0, 0.0, 0.0]), np.array([0.095, 0.0, 0.0])
    swap, ps, pm, n, g = oracle.sphere_contact(ca, 0.05, cb, 0.05)
    assert not swap                                     # a has the smaller centre: the slave
    np.testing.assert_allclose(n, [1, 0, 0], atol=1e-15)
    np.testing.assert_allclose(ps, [0.0475, 0, 0], atol=1e-15)   # overlap: points merged at the midpoint
    np.testing.assert_allclose(pm, ps, atol=0)
    assert g == pytest.approx(-0.005, abs=1e-15)
    # the same pair given the other way round: same slave, same normal
    swap2, ps2, pm2, n2, g2 = oracle.sphere_contact(cb, 0.05, ca, 0.05)
    assert swap2
    np.testing.assert_allclose(n2, n, atol=1e-15)
    assert g2 == pytest.approx(g, abs=1e-15)
    # apart by less than GEOMETRIC_EPSILON: a contact with zero gap (goc.c:50-59 clips at 0)
    swap, ps, pm, n, g = oracle.sphere_contact(ca, 0.05, [0.1000005, 0, 0], 0.05)
    assert g == 0.0
    np.testing.assert_allclose(ps, [0.05, 0, 0], atol=1e-15)
    np.testing.assert_allclose(pm, [0.0500005, 0, 0], atol=1e-15)
    # apart by more: none
    assert oracle.sphere_contact(ca, 0.05, [0.1000011, 0, 0], 0.05) is None
    # lexicographic order decides along y when x ties within the tolerance
    swap, *_ = oracle.sphere_contact([0.0, 0.05, 0.0], 0.05, [0.0000005, -0.04, 0.0], 0.05)
    assert swap


def test_plane_sphere_closed_forms():
    P = np.array([0.0, 0.0, 0.0, 0.0, 0.0, 1.0])
    pp, ps, n, g = oracle.plane_contact(P, [0.3, 0.2, 0.048], 0.05)
    np.testing.assert_allclose(pp, [0.3, 0.2, 0.0], atol=1e-15)
    np.testing.assert_allclose(ps, [0.3, 0.2, -0.002], atol=1e-15)
    np.testing.assert_allclose(n, [0, 0, 1], atol=0)
    assert g == pytest.approx(-0.002, abs=1e-15)
    assert oracle.plane_contact(P, [0.3, 0.2, 0.0500005], 0.05)[3] == 0.0
    assert oracle.plane_contact(P, [0.3, 0.2, 0.050002], 0.05) is None


@pytest.mark.parametrize("O", [[1e-4, -2e-4, 3e-5], [0.05, 0.1, -0.08], [0.3, -0.2, 0.5], [1.5, 0.7, -2.2]])
def test_expmap_against_rodrigues(O):
    np.testing.assert_allclose(oracle.expmap(O), rodrigues(np.array(O)), rtol=0, atol=2e-16 * (1 + np.linalg.norm(O)))


def test_free_fall_and_torque_free_spin():
    h, g = 1e-3, np.array([0.0, 0.0, -9.81])
    m, j = 2.0, 0.01
    J, I = (j * np.eye(3)).ravel(), (np.eye(3) / j).ravel()
    W0, v0 = np.array([3.0, -40.0, 25.0]), np.array([0.5, -0.2, 1.0])
    rot0, x0 = rodrigues(np.array([0.2, 0.1, -0.3])), np.array([1.0, 2.0, 3.0])
    velo, velo0, rot, x, aux = oracle.rig_step_begin(0, h, g, m, 1 / m, J, I, np.r_[W0, v0], rot0, x0)
    np.testing.assert_array_equal(velo0, np.r_[W0, v0])
    np.testing.assert_allclose(x, x0 + 0.5 * h * v0, rtol=1e-15)
    np.testing.assert_allclose(velo[3:], v0 + h * g, rtol=1e-15)
    np.testing.assert_allclose(velo[:3], W0, rtol=1e-14)        # isotropic body: W x J W = 0
    velo, rot, x, aux = oracle.rig_step_end(0, h, I, np.zeros(6), velo, rot, x, aux)
    np.testing.assert_allclose(x, x0 + 0.5 * h * (2 * v0 + h * g), rtol=1e-15)
    np.testing.assert_allclose(velo[:3], W0, rtol=1e-13)
    Rexp = (rot0.reshape(3, 3).T @ rodrigues(h * W0).reshape(3, 3).T).T.ravel()
    np.testing.assert_allclose(rot, Rexp, atol=1e-15)
    # obstacles do not move
    velo, velo0, rot, x, aux = oracle.rig_step_begin(1, h, g, m, 1 / m, J, I, np.zeros(6), rot0, x0)
    np.testing.assert_array_equal(x, x0)
    np.testing.assert_array_equal(velo, np.zeros(6))


# ---------------- host layer against the oracle ----------------

def _planes(L):
    return np.array([[0, 0, 0, 1, 0, 0], [L, 0, 0, -1, 0, 0], [0, 0, 0, 0, 1, 0], [0, L, 0, 0, -1, 0],
                     [0, 0, 0, 0, 0, 1], [0, 0, L, 0, 0, -1]], float)


def _geometry(n, L):
    shape = np.r_[np.zeros(n, np.int32), np.ones(6, np.int32)]
    plane = np.zeros((n + 6, 6))
    plane[n:] = _planes(L)
    return shape, plane


def _compare(s, O, n_old_R=None):
    c, b = s.contacts(), s.bodies()
    np.testing.assert_array_equal(c["master"], O["master"])
    np.testing.assert_array_equal(c["slave"], O["slave"])
    base = c["base"].reshape(-1, 3, 3)
    np.testing.assert_allclose(base[:, 2, :], O["normal"], rtol=0, atol=1e-14)
    np.testing.assert_allclose(c["gap"], O["gap"], rtol=0, atol=1e-15)
    # referential points: X = X0 + R'(p - x)
    x, x0, rot = b["x"].reshape(-1, 3), b["x0"].reshape(-1, 3), b["rot"].reshape(-1, 3, 3)
    for key, pts, bod in (("mpnt", O["mpnt"], O["master"]), ("spnt", O["spnt"], O["slave"])):
        ref = x0[bod] + np.einsum("kji,kj->ki", rot[bod].transpose(0, 2, 1), pts - x[bod])
        np.testing.assert_allclose(c[key].reshape(-1, 3), ref, rtol=0, atol=1e-14)
    return c


def test_host_detection_matches_brute_force():
    n, L = 343, 0.6           # spacing 0.0857 < 2 r: neighbours overlap
    s = lattice(n, L=L, seed=5, steps_begin=0)
    b0 = s.bodies()
    s.step_begin()
    shape, plane = _geometry(n, L)
    b = s.bodies()
    # the half step of every body against the oracle's
    mass = np.where(b0["minv"] > 0, 1.0 / np.where(b0["minv"] > 0, b0["minv"], 1.0), 0.0)
    for k in range(0, n + 6, 7):
        I = b0["jinv"][9 * k:9 * k + 9]
        J = np.linalg.inv(I.reshape(3, 3).T).T.ravel() if b0["kind"][k] == 0 else np.zeros(9)
        velo, velo0, rot, x, _ = oracle.rig_step_begin(b0["kind"][k], 1e-3, [0, 0, -9.81], mass[k], b0["minv"][k], J, I,
                                                        b0["velo"][6 * k:6 * k + 6], b0["rot"][9 * k:9 * k + 9],
                                                        b0["x"][3 * k:3 * k + 3])
        np.testing.assert_allclose(b["x"][3 * k:3 * k + 3], x, rtol=1e-15)
        np.testing.assert_allclose(b["velo"][6 * k:6 * k + 6], velo, rtol=1e-15, atol=1e-18)
        np.testing.assert_allclose(b["rot"][9 * k:9 * k + 9], rot, atol=1e-15)
    O = oracle.detect(shape, b["radius"], plane.ravel(), b["x"])
    assert len(O["master"]) > 2 * n
    c1 = _compare(s, O)
    # move the spheres apart and together at random: some contacts go, some come
    rng = np.random.default_rng(3)
    for k in range(n):
        s.set_velocity(k, np.r_[rng.normal(scale=50.0, size=3), rng.normal(scale=8.0, size=3)])
    s.step_begin()
    b = s.bodies()
    O = oracle.detect(shape, b["radius"], plane.ravel(), b["x"], c1["master"], c1["slave"])
    kept = (O["old"] >= 0).sum()
    assert 0 < kept < len(c1["master"]) and kept < len(O["master"])
    _compare(s, O)


# ---------------- QUASI_STATIC rigid step (bod.c:357-386, 1588-1645) ----------------

def _inertia():
    J = np.diag([0.01, 0.02, 0.035])
    return J.ravel(), np.linalg.inv(J).ravel()


def test_quasi_static_step_forgets_the_velocity():
    """u(t+h) = h M^-1 f(t+h): under gravity alone a body starts every step from
    rest, whatever it carried, and its velocity is h g."""
    h, g, m = 1e-3, np.array([0.0, 0.0, -9.81]), 2.0
    J, I = _inertia()
    velo = oracle.rig_static_begin(0, h, g, m, 1 / m, J, I, np.r_[0.0, 0.0, 0.0, 0.5, -0.2, 1.0])
    np.testing.assert_array_equal(velo[:3], 0.0)
    np.testing.assert_allclose(velo[3:], h * g, rtol=1e-15)
    # spin about a principal axis: exp[-hW] J W = J W, W1 = W, W1 x J W1 = 0: the spin stops
    velo = oracle.rig_static_begin(0, h, g, m, 1 / m, J, I, np.r_[0.0, 0.0, 25.0, 0.0, 0.0, 0.0])
    assert np.abs(velo[:3]).max() <= 1e-13 * 25.0
    # obstacles keep their (zero) velocity
    np.testing.assert_array_equal(oracle.rig_static_begin(1, h, g, m, 1 / m, J, I, np.zeros(6)), np.zeros(6))


def test_quasi_static_step_is_eulers_equation_as_h_goes_to_zero():
    """Off a principal axis the angular velocity after the step is h times the
    torque-free Euler acceleration -J^-1 (W x J W), up to O(h^2)."""
    g, m = np.zeros(3), 2.0
    J, I = _inertia()
    Jm = J.reshape(3, 3)
    W = np.array([3.0, -4.0, 5.0])
    euler = -np.linalg.solve(Jm, np.cross(W, Jm @ W))
    errs = []
    for h in (1e-3, 5e-4, 2.5e-4, 1.25e-4):
        velo = oracle.rig_static_begin(0, h, g, m, 1 / m, J, I, np.r_[W, 0.0, 0.0, 0.0])
        errs.append(np.abs(velo[:3] / h - euler).max())
    print("Euler residuals", errs, "of", np.abs(euler).max())
    assert errs[0] < 1e-2 * np.abs(euler).max()
    for a, b in zip(errs, errs[1:]):
        assert 1.9 < a / b < 2.1          # first order in h


def test_quasi_static_end_step():
    """x(t+h) = x(t) + h v, R(t+h) = R(t) exp[h W] (a Rodrigues rotation)."""
    h = 1e-3
    velo = np.r_[3.0, -40.0, 25.0, 0.5, -0.2, 1.0]
    rot0, x0 = rodrigues(np.array([0.2, 0.1, -0.3])), np.array([1.0, 2.0, 3.0])
    rot, x = oracle.rig_static_end(0, h, velo, rot0, x0)
    np.testing.assert_allclose(x, x0 + h * velo[3:], rtol=1e-15)
    Rexp = (rot0.reshape(3, 3).T @ rodrigues(h * velo[:3]).reshape(3, 3).T).T.ravel()
    np.testing.assert_allclose(rot, Rexp, atol=1e-15)
    rot, x = oracle.rig_static_end(1, h, velo, rot0, x0)
    np.testing.assert_array_equal(x, x0)


def test_host_quasi_static_step_matches_oracle():
    """The host layer's QUASI_STATIC step_begin (shared with the device loop)
    against the oracle, body by body; the previous velocities stay."""
    n = 343
    s = lattice(n, L=0.6, seed=5, steps_begin=0, analysis="QUASI_STATIC")
    rng = np.random.default_rng(2)
    for k in range(n):
        s.set_velocity(k, np.r_[rng.normal(scale=5.0, size=3), rng.normal(scale=0.5, size=3)])
    b0 = s.bodies()
    s.step_begin()
    b = s.bodies()
    np.testing.assert_array_equal(b["velo0"], b0["velo0"])
    np.testing.assert_array_equal(b["x"], b0["x"])
    np.testing.assert_array_equal(b["rot"], b0["rot"])
    mass = np.where(b0["minv"] > 0, 1.0 / np.where(b0["minv"] > 0, b0["minv"], 1.0), 0.0)
    for k in range(n + 6):
        I = b0["jinv"][9 * k:9 * k + 9]
        J = np.linalg.inv(I.reshape(3, 3).T).T.ravel() if b0["kind"][k] == 0 else np.zeros(9)
        velo = oracle.rig_static_begin(b0["kind"][k], 1e-3, [0, 0, -9.81], mass[k], b0["minv"][k], J, I,
                                       b0["velo"][6 * k:6 * k + 6])
        # J here is inv(jinv), not the host's own: the spheres' gyroscopic terms
        # (zero in exact arithmetic) differ at the rounding level
        np.testing.assert_allclose(b["velo"][6 * k:6 * k + 6], velo, rtol=1e-13, atol=1e-16)
