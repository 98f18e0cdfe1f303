"""Pins the oracle's cone operations (or_scf.c, restating scf.c:28-174) against
the mathematics of the Coulomb cone K = {(t, n) : |t| <= mu n}: the textbook
projection onto a second-order cone, its idempotence and polar-cone residual,
the zeros of the contact residual C(U, R) at exact stick / slide / separation
states, and the Jacobians X, Y against central differences."""
import ctypes as C

import numpy as np
import pytest

import oracle

L = oracle.lib()


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def project(mu, coh, S):
    S = np.array(S, float)
    R = np.zeros(3)
    L.OR_SCF_Project(mu, coh, _p(S), _p(R))
    return R


def soc_textbook(mu, x):
    """Projection onto {(t, n): |t| <= mu n} (closed form for a circular cone)."""
    t, n = x[:2], x[2]
    nt = np.linalg.norm(t)
    if nt <= mu * n:
        return x.copy()
    if mu * nt <= -n:
        return np.zeros(3)
    a = (mu * nt + n) / (1.0 + mu * mu)
    return np.array([mu * a * t[0] / nt, mu * a * t[1] / nt, a])


def residual(U, R, mu, res=0.0, coh=0.0, V=(0, 0, 0), omega=0.0, UT=-1.0, want_jac=False):
    U, R, V = np.array(U, float), np.array(R, float), np.array(V, float)
    Cv, X, Y = np.zeros(3), np.zeros(9), np.zeros(9)
    L.OR_SCF_Linearize(1, 1e-3, mu, res, coh, 0.0, _p(V), _p(U), _p(R), UT, omega, _p(Cv),
                       _p(X) if want_jac else None, _p(Y) if want_jac else None)
    if want_jac:
        return Cv, X.reshape(3, 3).T, Y.reshape(3, 3).T
    return Cv


@pytest.mark.parametrize("mu", [0.1, 0.3, 0.8, 1.5])
def test_projection_matches_textbook_formula(mu):
    rng = np.random.default_rng(1)
    for _ in range(500):
        x = rng.normal(size=3) * rng.choice([1e-3, 1.0, 1e3])
        assert np.allclose(project(mu, 0.0, x), soc_textbook(mu, x), rtol=1e-12, atol=1e-12 * np.abs(x).max())


def test_projection_properties():
    rng = np.random.default_rng(2)
    for _ in range(300):
        mu = rng.uniform(0.05, 1.2)
        x = rng.normal(size=3)
        p = project(mu, 0.0, x)
        # membership, idempotence, polar residual, orthogonality
        assert np.linalg.norm(p[:2]) <= mu * p[2] * (1 + 1e-12) + 1e-14
        assert np.allclose(project(mu, 0.0, p), p, atol=1e-13)
        q = x - p
        assert mu * np.linalg.norm(q[:2]) <= -q[2] * (1 + 1e-12) + 1e-13
        assert abs(np.dot(p, q)) <= 1e-12 * max(1.0, np.dot(x, x))


def test_projection_cohesion_shift():
    rng = np.random.default_rng(3)
    for _ in range(100):
        mu, c = rng.uniform(0.1, 1.0), rng.uniform(0.0, 2.0)
        x = rng.normal(size=3)
        shifted = x + np.array([0.0, 0.0, c])
        assert np.allclose(project(mu, c, x), soc_textbook(mu, shifted) - np.array([0.0, 0.0, c]), atol=1e-13)


def test_residual_zero_at_sticking():
    rng = np.random.default_rng(4)
    for _ in range(100):
        mu = rng.uniform(0.1, 1.0)
        rn = rng.uniform(0.1, 5.0)
        rt = rng.normal(size=2)
        rt *= 0.9 * mu * rn * rng.uniform(0, 1) / max(np.linalg.norm(rt), 1e-300)
        Cv = residual([0.0, 0.0, 0.0], [rt[0], rt[1], rn], mu)
        assert np.abs(Cv).max() <= 1e-13 * rn


def test_residual_zero_at_separation():
    rng = np.random.default_rng(5)
    for _ in range(100):
        mu = rng.uniform(0.1, 1.0)
        U = np.array([*rng.normal(size=2), rng.uniform(0.0, 3.0)])
        Cv = residual(U, [0.0, 0.0, 0.0], mu)
        assert np.abs(Cv).max() <= 1e-13 * max(1.0, np.abs(U).max())


def test_residual_zero_at_sliding():
    rng = np.random.default_rng(6)
    for _ in range(100):
        mu = rng.uniform(0.1, 1.0)
        ut = rng.normal(size=2)
        rn = rng.uniform(0.1, 4.0)
        rt = -mu * rn * ut / np.linalg.norm(ut)
        Cv = residual([ut[0], ut[1], 0.0], [rt[0], rt[1], rn], mu)
        assert np.abs(Cv).max() <= 1e-12 * max(rn, np.abs(ut).max())


def test_residual_with_zero_reaction_is_minus_projection():
    """C(U, 0) = F + m(-F) = -proj_K(-F(U)), F = [U_T, U_N + mu |U_T|]."""
    rng = np.random.default_rng(7)
    for _ in range(200):
        mu = rng.uniform(0.1, 1.0)
        U = rng.normal(size=3)
        F = np.array([U[0], U[1], U[2] + mu * np.linalg.norm(U[:2])])
        assert np.allclose(residual(U, [0.0, 0.0, 0.0], mu), -soc_textbook(mu, -F), atol=1e-13)


def test_restitution_enters_the_normal_velocity():
    U = np.array([0.1, -0.2, 0.3])
    V = np.array([0.0, 0.0, -2.0])
    e = 0.5
    c1 = residual(U, [0.0, 0.0, 0.0], 0.3, res=e, V=V)
    U2 = U + np.array([0.0, 0.0, e * min(V[2], 0.0)])
    c2 = residual(U2, [0.0, 0.0, 0.0], 0.3, res=0.0)
    assert np.allclose(c1, c2, atol=1e-15)


def test_jacobians_match_finite_differences():
    rng = np.random.default_rng(8)
    h = 1e-6
    for _ in range(50):
        mu, om = rng.uniform(0.1, 1.0), rng.uniform(0.05, 0.5)
        U, R = rng.normal(size=3), rng.normal(size=3)
        _, X, Y = residual(U, R, mu, omega=om, want_jac=True)
        Xf, Yf = np.zeros((3, 3)), np.zeros((3, 3))
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            Xf[:, k] = (residual(U + e, R, mu, omega=om) - residual(U - e, R, mu, omega=om)) / (2 * h)
            Yf[:, k] = (residual(U, R + e, mu, omega=om) - residual(U, R - e, mu, omega=om)) / (2 * h)
        assert np.allclose(X, Xf, atol=1e-6), (X, Xf)
        assert np.allclose(Y, Yf, atol=1e-6), (Y, Yf)
