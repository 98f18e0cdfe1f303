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


def test_cohesive_contact_holds_tension_down_to_its_cohesion():
    W = np.diag([2.0, 2.0, 1.5])
    # a positive gap and a separating normal velocity: the dashpot pulls, R_N = -c V_N = -1
    B, V = np.array([0.0, 0.0, 0.2]), np.array([0.0, 0.0, 0.1])
    # cohesion 5 > 1: the tension is kept, the contact stays cohesive, R_T cancels B_T (none here)
    rc, R, U, coh = oracle.penalty_contact(0, H, 1e-3, 1e5, 10.0, 1.0, 0.5, cm(W), B, V, cohesion=5.0)
    assert rc == 0 and coh == 5.0
    np.testing.assert_array_equal(R, [0.0, 0.0, -1.0])
    np.testing.assert_allclose(U, B + W @ R, rtol=1e-15)
    # cohesion 0.4 < 1: mode I, R_N is held at -0.4 and the cohesion is lost
    rc, R, U, coh = oracle.penalty_contact(0, H, 1e-3, 1e5, 10.0, 1.0, 0.5, cm(W), B, V, cohesion=0.4)
    assert coh == 0.0
    np.testing.assert_array_equal(R, [0.0, 0.0, -0.4])
    # without cohesion the same contact separates
    rc, R, U, coh = oracle.penalty_contact(0, H, 1e-3, 1e5, 10.0, 1.0, 0.5, cm(W), B, V, cohesion=0.0)
    np.testing.assert_array_equal(R, 0.0)
    np.testing.assert_array_equal(U, 0.0)


def test_cohesive_contact_breaks_on_the_cone():
    W = np.diag([2.0, 2.0, 1.5])
    B, V = np.array([3.0, 4.0, -0.1]), np.array([0.0, 0.0, 0.0])
    # R_N = k (-g) = 100, unclamped |R_T| = 2.5 <= mu R_N = 30: sticks, cohesion kept
    rc, R, U, coh = oracle.penalty_contact(0, H, -1e-3, 1e5, 0.0, 1.0, 0.3, cm(W), B, V, cohesion=2.0)
    assert coh == 2.0
    np.testing.assert_allclose(R[:2], [-1.5, -2.0], rtol=1e-15)
    # |R_T| = 250 > 30: clamped to the cone (mode II) and the cohesion is lost
    rc, R, U, coh = oracle.penalty_contact(0, H, -1e-3, 1e5, 0.0, 1.0, 0.3, cm(W), 100 * B, V, cohesion=2.0)
    assert coh == 0.0
    np.testing.assert_allclose(np.hypot(R[0], R[1]), 30.0, rtol=1e-14)


def test_update_cohesion_modes_on_either_side_of_their_bounds():
    c, mu = 2.0, 0.5
    e = 1e-8 * c
    R = np.array([
        [0.0, 0.0, -c + 4 * e],               # R_N + c = 4e >= e, |R_T| + e = e < mu 4e = 2e: kept
        [0.0, 0.0, -c + 2 * e],               # R_N + c = 2e >= e, but |R_T| + e = e >= mu 2e: mode II
        [0.0, 0.0, -c + 0.5 * e],             # R_N + c < e: mode I
        [0.3, 0.4, 1.0],                      # |R_T| + e = 0.5 + e < mu (1 + c) = 1.5: kept
        [0.9, 1.2 - 1e-6, 1.0],               # |R_T| + e just below 1.5: kept
        [0.9, 1.2, 1.0],                      # |R_T| = 1.5, + e >= 1.5: mode II
        [0.0, 0.0, -10.0],                    # not cohesive: left at 0
    ])
    coh = np.array([c, c, c, c, c, c, 0.0])
    out = oracle.update_cohesion(R, np.full(7, mu), coh)
    np.testing.assert_array_equal(out, [c, 0.0, 0.0, c, c, 0.0, 0.0])
