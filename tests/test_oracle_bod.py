"""Pins the oracle body update (or_bod.c, restating bod.c:418-435, 1446-1447):
equal and opposite linear impulses on an isolated pair, no torque when the
reaction acts along the normal at a sphere's surface point, the torque
(x_p - x_c) x f of a tangential reaction, and obstacles untouched."""
import numpy as np

import oracle
import solfec

H = 1e-3


def pair():
    s = solfec.SOLFEC("DYNAMIC", H, "/tmp/solfec_test")
    solfec.SURFACE_MATERIAL(s, friction=0.3)
    bulk = solfec.BULK_MATERIAL(s, density=2000.0)
    solfec.BODY(s, "RIGID", solfec.SPHERE((0.0, 0.0, 0.0), 0.05, 1, 1), bulk)
    solfec.BODY(s, "RIGID", solfec.SPHERE((0.0, 0.0, 0.0999), 0.05, 1, 1), bulk)
    solfec.BODY(s, "OBSTACLE", solfec.PLANE((0, 0, -0.0499), (0, 0, 1), 0), bulk)
    s.step_begin()
    return s.bodies(), s.contacts()


def test_pair_momentum_balance_and_torques():
    b, c = pair()
    n = len(c["master"])
    assert n == 2
    bo, co = oracle.bodies_from(b), oracle.contacts_from(c)
    m = 2000.0 * 4.0 / 3.0 * np.pi * 0.05 ** 3
    # normal reaction only on the sphere-sphere contact
    ss = [i for i in range(n) if c["slave"][i] != 2][0]
    R = np.zeros(3 * n)
    R[3 * ss + 2] = 5.0
    velo, force = oracle.body_update(bo, co, R, H)
    assert np.allclose(force[0, 3:] + force[1, 3:], 0.0, atol=1e-12)
    assert np.allclose(force[:2, :3], 0.0, atol=1e-10)          # along the normal: no torque
    assert np.allclose(force[2], 0.0)                           # the obstacle takes nothing
    dv = velo[:2, 3:] - b["velo"].reshape(-1, 6)[:2, 3:]
    assert np.allclose(dv[0] + dv[1], 0.0, atol=1e-12)
    assert abs(np.linalg.norm(dv[0]) - H * 5.0 / m) < 1e-12
    # tangential reaction: torque = (x_p - x_c) x f about each centre
    R = np.zeros(3 * n)
    R[3 * ss + 0] = 2.0
    velo, force = oracle.body_update(bo, co, R, H)
    base = c["base"][9 * ss:9 * ss + 9].reshape(3, 3).T
    f = base @ np.array([2.0, 0.0, 0.0])
    x = b["x"].reshape(-1, 3)
    m_id, s_id = c["master"][ss], c["slave"][ss]
    for body, sign, pnt in ((m_id, 1.0, c["mpnt"][3 * ss:3 * ss + 3]), (s_id, -1.0, c["spnt"][3 * ss:3 * ss + 3])):
        xp = x[body] + (pnt - b["x0"][3 * body:3 * body + 3])    # R = I at this point of the run
        assert np.allclose(force[body, :3], np.cross(xp - x[body], sign * f), atol=1e-12)
        assert np.allclose(force[body, 3:], sign * f, atol=1e-12)
