"""The spring-dashpot contact law of the penalty solver (oracle or_pes.c,
pes.c:33-106) against closed forms."""
import numpy as np
import pytest

import oracle

H = 1e-3


def cm(M):
    return np.asarray(M, float).T.ravel()     # column-major 9-vector


def test_explicit_normal_reaction_and_sticking():
    W = np.diag([2.0, 3.0, 1.5])
    B, V = np.array([0.02, -0.01, -0.3]), np.array([0.0, 0.0, -0.2])
    k, c, g = 1e5, 10.0, -1e-3
    rc, R, U = oracle.penalty_contact(0, H, g, k, c, 1.0, 0.9, cm(W), B, V)
    assert rc == 0
    RN = k * 1e-3 - c * (-0.2)                 # k (-g) - c V_N
    assert R[2] == pytest.approx(RN, rel=1e-15)
    # tangential reaction cancels the tangential velocity (diagonal W: R_T = -B_T / W_TT), inside the cone
    np.testing.assert_allclose(R[:2], [-0.02 / 2.0, 0.01 / 3.0], rtol=1e-15)
    np.testing.assert_allclose(U, B + W @ R, rtol=1e-14, atol=1e-17)
    np.testing.assert_allclose(U[:2], 0.0, atol=1e-17)


def test_implicit_normal_reaction_and_hertz_power():
    W = cm(np.diag([2.0, 2.0, 1.5]))
    B, V = np.array([0.0, 0.0, -0.3]), np.array([0.0, 0.0, -0.2])
    k, c, p, g = 2e6, 5.0, 1.5, -2e-4
    rc, R, U = oracle.penalty_contact(1, H, g, k, c, p, 0.5, W, B, V)
    ge = g + 0.25 * H * (B[2] - V[2])          # gap a quarter step ahead
    s = k * p * (-ge) ** (p - 1.0)
    RN = (k * (-ge) ** p - 0.5 * c * (B[2] + V[2])) / (1.0 + (0.25 * H * s + 0.5 * c) * 1.5)
    assert R[2] == pytest.approx(RN, rel=1e-14)
    # critical damping 2 sqrt(h s / W_NN)
    rc, R2, _ = oracle.penalty_contact(1, H, g, k, -1.0, p, 0.5, W, B, V)
    cc = 2.0 * np.sqrt(H * s / 1.5)
    RN2 = (k * (-ge) ** p - 0.5 * cc * (B[2] + V[2])) / (1.0 + (0.25 * H * s + 0.5 * cc) * 1.5)
    assert R2[2] == pytest.approx(RN2, rel=1e-14)


def test_separation_and_sliding():
    W = cm(np.diag([2.0, 2.0, 1.5]))
    # a positive gap and a separating normal velocity: no spring force, the dashpot pulls: R_N < 0 -> 0
    rc, R, U = oracle.penalty_contact(0, H, 1e-3, 1e5, 10.0, 1.0, 0.5, W, [0.1, 0, 0.2], [0, 0, 0.1])
    np.testing.assert_array_equal(R, 0.0)
    np.testing.assert_array_equal(U, 0.0)
    # a large tangential velocity: the tangential reaction is clamped to mu R_N, along -B_T
    B = np.array([300.0, 400.0, -0.1])        # unclamped |R_T| = 250 > mu R_N = 30
    rc, R, U = oracle.penalty_contact(0, H, -1e-3, 1e5, 0.0, 1.0, 0.3, W, B, [0, 0, 0])
    RN = 1e5 * 1e-3
    assert R[2] == pytest.approx(RN)
    np.testing.assert_allclose(np.hypot(R[0], R[1]), 0.3 * RN, rtol=1e-14)
    np.testing.assert_allclose(R[:2] / np.hypot(R[0], R[1]), -B[:2] / 500.0, rtol=1e-14)


def test_carried_tangential_reaction_enters_the_normal_velocity():
    W = np.array([[2.0, 0.0, 0.3], [0.0, 2.0, 0.1], [0.3, 0.1, 1.5]])
    B, V = np.array([0.0, 0.0, -0.3]), np.array([0.0, 0.0, -0.2])
    R0 = np.array([0.5, -0.2, 0.0])
    rc, R, U = oracle.penalty_contact(1, H, -1e-4, 1e5, 3.0, 1.0, 0.5, cm(W), B, V, R0)
    BN = B[2] + W[2, 0] * R0[0] + W[2, 1] * R0[1]
    ge = -1e-4 + 0.25 * H * (BN - V[2])
    RN = (1e5 * (-ge) - 0.5 * 3.0 * (BN + V[2])) / (1.0 + (0.25 * H * 1e5 + 0.5 * 3.0) * 1.5)
    assert R[2] == pytest.approx(RN, rel=1e-14)


def test_singular_tangential_block_is_an_error():
    W = cm([[1.0, 1.0, 0.0], [1.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    rc, R, U = oracle.penalty_contact(0, H, -1e-3, 1e5, 0.0, 1.0, 0.3, W, [0.1, 0, 0], [0, 0, 0])
    assert rc == -1
