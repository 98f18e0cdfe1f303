"""The Python surface keeps the reference's GAUSS_SEIDEL_SOLVER / RUN contract
(lng.c:4173-4641, 8973-9025): keywords, defaults, derived diagepsilon and
diagmaxiter, read-only and read/write members, validation errors.  The host
domain (scene, contact detection) is checked against brute force.  No GPU."""
import numpy as np
import pytest

import solfec
from scenes import lattice


def test_defaults_and_members():
    gs = solfec.GAUSS_SEIDEL_SOLVER(1e-4, 1000)
    assert gs.epsilon == 1e-4 and gs.maxiter == 1000 and gs.meritval == 1.0
    assert gs.failure == "FAILURE_CONTINUE" and gs.diagsolver == "SEMISMOOTH_NEWTON"
    # lng.c:4336-4342: 0.01 min(eps, meritval, 1e-4), max(100, maxiter/100)
    assert gs.diagepsilon == pytest.approx(1e-6) and gs.diagmaxiter == 100
    gs2 = solfec.GAUSS_SEIDEL_SOLVER(1e-6, 50000, meritval=1e-8)
    assert gs2.diagepsilon == pytest.approx(1e-10) and gs2.diagmaxiter == 500
    gs3 = solfec.GAUSS_SEIDEL_SOLVER(1e-20, 10)
    assert gs3.diagepsilon == pytest.approx(np.finfo(float).eps)
    assert gs.reverse == "OFF" and gs.variant == "FULL" and gs.innerloops == 1
    assert gs.iters == 0 and gs.error == "OK" and gs.rerhist == [] and gs.itershist is None


def test_read_write_members():
    gs = solfec.GAUSS_SEIDEL_SOLVER(1e-4, 1000)
    gs.epsilon = 1e-5
    gs.maxiter = 20
    gs.diagepsilon = 1e-9
    gs.diagmaxiter = 7
    gs.diagsolver = "PROJECTED_GRADIENT"
    gs.reverse = "ON"
    gs.variant = "MIDDLE_JACOBI"
    gs.innerloops = 3
    gs.precision = "MIXED"
    assert (gs.epsilon, gs.maxiter, gs.diagepsilon, gs.diagmaxiter) == (1e-5, 20, 1e-9, 7)
    assert gs.diagsolver == "PROJECTED_GRADIENT" and gs.reverse == "ON" and gs.variant == "MIDDLE_JACOBI"
    assert gs.innerloops == 3 and gs.precision == "MIXED"
    for member in ("failure", "error", "iters", "rerhist", "merhist"):
        with pytest.raises(ValueError):
            setattr(gs, member, 1)
    with pytest.raises(ValueError):
        gs.diagsolver = "NEWTON"
    with pytest.raises(ValueError):
        gs.reverse = "MAYBE"
    with pytest.raises(ValueError):
        gs.epsilon = -1.0


def test_constructor_validation():
    with pytest.raises(ValueError):
        solfec.GAUSS_SEIDEL_SOLVER(-1.0, 10)
    with pytest.raises(ValueError):
        solfec.GAUSS_SEIDEL_SOLVER(1e-4, 0)
    with pytest.raises(ValueError):
        solfec.GAUSS_SEIDEL_SOLVER(1e-4, 10, failure="ABORT")
    with pytest.raises(ValueError):
        solfec.GAUSS_SEIDEL_SOLVER(1e-4, 10, failure="CALLBACK")
    with pytest.raises(ValueError):
        solfec.GAUSS_SEIDEL_SOLVER(1e-4, 10, diagsolver="X")
    gs = solfec.GAUSS_SEIDEL_SOLVER(1e-4, 10, failure="CALLBACK", callback=lambda s: 1, data=(1, 2))
    assert gs.failure == "FAILURE_CALLBACK"
    with pytest.raises(TypeError):
        solfec.RUN(object(), gs, 1.0)


def test_scene_is_seeded_and_contacts_match_brute_force():
    s1 = lattice(343, seed=5)
    s2 = lattice(343, seed=5)
    b1, c1 = s1.bodies(), s1.contacts()
    b2, c2 = s2.bodies(), s2.contacts()
    assert np.array_equal(b1["x"], b2["x"]) and np.array_equal(c1["master"], c2["master"])
    x = b1["x"].reshape(-1, 3)
    r = b1["radius"]
    sph = np.where(b1["kind"] == 0)[0]
    want = set()
    for a in sph:
        d = np.linalg.norm(x[sph] - x[a], axis=1) - (r[sph] + r[a])
        for bb in sph[(d < 1e-6) & (sph > a)]:
            want.add((int(a), int(bb)))
    got = {(min(m, s), max(m, s)) for m, s in zip(c1["master"], c1["slave"]) if b1["kind"][s] == 0}
    assert got == want
    # normal from slave to master, unit, gap <= 0, frame orthonormal
    n = len(c1["master"])
    for i in range(n):
        base = c1["base"][9 * i:9 * i + 9].reshape(3, 3).T
        assert np.allclose(base.T @ base, np.eye(3), atol=1e-12)
        assert c1["gap"][i] <= 0.0
        m, s = c1["master"][i], c1["slave"][i]
        if b1["kind"][s] == 0:
            d = x[m] - x[s]
            assert np.allclose(base[:, 2], d / np.linalg.norm(d), atol=1e-12)


def test_contacts_persist_with_their_reactions():
    s = lattice(125, seed=2)
    c0 = s.contacts()
    n = len(c0["master"])
    s.set_RU(np.arange(3 * n, dtype=float), None)
    s.step_end()
    s.step_begin()
    c1 = s.contacts()
    keep = {(m, sl): k for k, (m, sl) in enumerate(zip(c0["master"], c0["slave"]))}
    for i, (m, sl) in enumerate(zip(c1["master"], c1["slave"])):
        if (m, sl) in keep:
            k = keep[(m, sl)]
            assert np.array_equal(c1["R"][3 * i:3 * i + 3], np.arange(3 * k, 3 * k + 3, dtype=float))
        else:
            assert np.all(c1["R"][3 * i:3 * i + 3] == 0.0)


def _domain():
    s = solfec.SOLFEC("DYNAMIC", 1e-3, "/tmp/solfec_test")
    solfec.SURFACE_MATERIAL(s, friction=0.3, restitution=0.0)
    return s


def _rot(axis, angle):
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * K @ K


def test_ellipsoid_mass_characteristics():
    s = _domain()
    R = _rot((1, 2, 3), 0.7)
    mat = solfec.BULK_MATERIAL(s, density=2000.0)
    solfec.BODY(s, "RIGID", solfec.ELLIP((0.1, 0.2, 0.3), (0.05, 0.04, 0.03), 1, 1, rotation=R.T.ravel()), mat)
    b = s.bodies()
    a, bb, c = 0.05, 0.04, 0.03
    m = 2000.0 * 4.0 / 3.0 * np.pi * a * bb * c
    J = R @ np.diag([m / 5 * (bb * bb + c * c), m / 5 * (a * a + c * c), m / 5 * (a * a + bb * bb)]) @ R.T
    np.testing.assert_allclose(b["minv"][0], 1.0 / m, rtol=1e-14)
    np.testing.assert_allclose(b["jinv"].reshape(3, 3).T, np.linalg.inv(J), rtol=1e-12)


def test_round_ellipsoids_contact_like_spheres():
    """An ellipsoid with equal semi-axes goes through the ellipsoid contact routine and
    must give the sphere routine's contact: point, frame and gap."""
    out = []
    for shape in ("sphere", "ellip"):
        s = _domain()
        mat = solfec.BULK_MATERIAL(s, density=2000.0)
        for x, r in (((0.0, 0.0, 0.0), 0.05), ((0.06, 0.07, 0.02), 0.06)):
            sh = solfec.SPHERE(x, r, 1, 1) if shape == "sphere" else solfec.ELLIP(x, (r, r, r), 1, 1)
            solfec.BODY(s, "RIGID", sh, mat)
        s.step_begin()
        out.append(s.contacts())
    a, b = out
    assert len(a["gap"]) == len(b["gap"]) == 1
    for k in ("master", "slave"):
        assert np.array_equal(a[k], b[k])
    for k in ("mpnt", "spnt", "base", "gap"):
        np.testing.assert_allclose(b[k], a[k], rtol=1e-9, atol=1e-12)


def test_ellipsoid_pair_and_plane_closed_forms():
    s = _domain()
    mat = solfec.BULK_MATERIAL(s, density=2000.0)
    a1, a2, dlt = 0.05, 0.03, 0.002
    # two axis-aligned ellipsoids on the x axis, overlapping by dlt along x
    solfec.BODY(s, "RIGID", solfec.ELLIP((0.5, 0.5, 0.5), (a1, 0.02, 0.04), 1, 1), mat)
    solfec.BODY(s, "RIGID", solfec.ELLIP((0.5 + a1 + a2 - dlt, 0.5, 0.5), (a2, 0.035, 0.025), 1, 1), mat)
    # a rotated ellipsoid above the plane z = 0, penetrating it by 0.001
    R = _rot((1, 1, 0), 0.9)
    A = R @ np.diag([0.06, 0.03, 0.02]) ** 2 @ R.T
    hz = np.sqrt(A[2, 2])
    solfec.BODY(s, "RIGID", solfec.ELLIP((2.0, 2.0, hz - 0.001), (0.06, 0.03, 0.02), 1, 1, rotation=R.T.ravel()), mat)
    solfec.BODY(s, "OBSTACLE", solfec.PLANE((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 1), mat)
    s.step_begin()
    c = s.contacts()
    assert len(c["gap"]) == 2
    n = len(c["gap"])
    base = c["base"].reshape(n, 3, 3)
    i = int(np.where(c["slave"] == 3)[0][0]) if np.any(c["slave"] == 3) else None
    j = 1 - i
    # pair: normal along +x from slave (first, smaller x centre) to master; gap = -dlt
    np.testing.assert_allclose(abs(base[j][2][0]), 1.0, rtol=1e-10)
    np.testing.assert_allclose(c["gap"][j], -dlt, rtol=1e-8)
    # plane: normal +z, gap -0.001, contact point at the ellipsoid's lowest point
    np.testing.assert_allclose(base[i][2], [0.0, 0.0, 1.0], atol=1e-14)
    np.testing.assert_allclose(c["gap"][i], -0.001, rtol=1e-9)
    x = s.bodies()["x"].reshape(-1, 3)[2]
    low = x - A @ np.array([0.0, 0.0, 1.0]) / hz
    X0 = s.bodies()["x0"].reshape(-1, 3)[2]
    np.testing.assert_allclose(c["mpnt"].reshape(n, 3)[i], X0 + (low - x), atol=1e-12)


def test_spring_dashpot_material_and_penalty_solver_validation():
    s = _domain()
    with pytest.raises(ValueError):
        solfec.SURFACE_MATERIAL(s, model="HERTZ")
    with pytest.raises(ValueError):
        solfec.SURFACE_MATERIAL(s, restitution=1.5)
    with pytest.raises(ValueError):
        solfec.SURFACE_MATERIAL(s, model="SPRING_DASHPOT", spring=-1.0)
    with pytest.raises(ValueError):
        solfec.SURFACE_MATERIAL(s, model="SPRING_DASHPOT", hpow=-0.5)
    m = solfec.SURFACE_MATERIAL(s, model="SPRING_DASHPOT", spring=1e6, dashpot=-1.0, hpow=1.5)
    assert (m.model, m.spring, m.dashpot, m.hpow) == ("SPRING_DASHPOT", 1e6, -1.0, 1.5)
    assert solfec.PENALTY_SOLVER().variant == "IMPLICIT"
    assert solfec.PENALTY_SOLVER("EXPLICIT").variant == "EXPLICIT"
    with pytest.raises(ValueError):
        solfec.PENALTY_SOLVER("SEMI")
    with pytest.raises(TypeError):
        solfec.RUN(s, object(), 0.001)


def test_gauss_seidel_refuses_spring_dashpot_materials():
    """The Gauss-Seidel sweep models SIGNORINI_COULOMB contacts only: a step is
    refused (before any device work) when the domain has a SPRING_DASHPOT pair."""
    s = _domain()
    solfec.SURFACE_MATERIAL(s, 0, 1, model="SPRING_DASHPOT", spring=1e6)
    solfec.LATTICE(s, 27, 0.3, 0.045, 0.055, 0.005, 2000.0, 1)
    assert s.step_begin() == 0
    assert s.step_solve(solfec.GAUSS_SEIDEL_SOLVER(1e-4, 100)) == 6
    assert "PENALTY_SOLVER" in s.lib.SOLFEC_Last_Error(s.sol).decode()
    with pytest.raises(RuntimeError, match="PENALTY_SOLVER"):
        solfec.RUN(s, solfec.GAUSS_SEIDEL_SOLVER(1e-4, 100), 0.001)


def test_newton_solver_members_and_validation():
    ns = solfec.NEWTON_SOLVER()
    assert (ns.meritval, ns.maxiter, ns.locdyn, ns.linver, ns.linmaxiter, ns.maxmatvec) == \
        (1e-8, 1000, "ON", "GMRES", 10, 10000)
    assert (ns.epsilon, ns.delta, ns.theta, ns.omega, ns.gsflag, ns.iters) == (0.25, 0.0, 0.25, 1e-10, "ON", 0)
    assert ns.merhist == [] and ns.mvhist == [] and ns.itershist is None
    ns.linver = "DIAG"
    ns.theta = 1.0
    ns.gsflag = "OFF"
    assert (ns.linver, ns.theta, ns.gsflag) == ("DIAG", 1.0, "OFF")
    ns.locdyn = "OFF"
    assert ns.locdyn == "OFF" and solfec.NEWTON_SOLVER(locdyn="OFF").locdyn == "OFF"
    for k, v in (("linver", "CG"), ("theta", 0.0), ("theta", 1.5), ("delta", -1.0), ("maxiter", 0),
                 ("locdyn", "MAYBE"), ("gsflag", "YES")):
        with pytest.raises(ValueError):
            setattr(ns, k, v)
    with pytest.raises(ValueError):
        ns.iters = 3
    for kw in (dict(meritval=0.0), dict(theta=2.0), dict(linver="LU"), dict(reldelta="meanWii"), dict(omega=0.0)):
        with pytest.raises(ValueError):
            solfec.NEWTON_SOLVER(**kw)
    assert solfec.NEWTON_SOLVER(reldelta="maxWii", linver="DIAG").linver == "DIAG"
