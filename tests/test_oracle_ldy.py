"""Pins the oracle's LOCDYN assembly (or_ldy.c, restating ldy.c:364-844 and
bod.c:216-258) against mechanics computed a different way: local velocities as
base' (v + omega x (x_p - x_c)) of the material points, W_ij as the local
velocity response at contact i to a unit local impulse at contact j over one
step, the sphere closed forms W_ii = (h/m) diag(7/2, 7/2, 1) (sphere on a
plane) and h diag(7/2 (1/m1+1/m2), ..., 1/m1+1/m2) (touching pair), numpy
eigvalsh / inv for rho and A, symmetry and positive semi-definiteness; the
impulse-response checks run on spheres and on rigid ellipsoids (general
inertia, normals off the centre lines)."""
import numpy as np
import pytest

import oracle
import solfec
from scenes import lattice

H = 1e-3


def body_arrays(b):
    nb = len(b["kind"])
    return dict(kind=b["kind"], R=b["rot"].reshape(nb, 3, 3).transpose(0, 2, 1), x=b["x"].reshape(nb, 3),
                x0=b["x0"].reshape(nb, 3), jinv=b["jinv"].reshape(nb, 3, 3).transpose(0, 2, 1),
                minv=b["minv"], velo=b["velo"].reshape(nb, 6), velo0=b["velo0"].reshape(nb, 6))


def contact_arrays(c):
    n = len(c["master"])
    return dict(master=c["master"], slave=c["slave"], mpnt=c["mpnt"].reshape(n, 3), spnt=c["spnt"].reshape(n, 3),
                base=c["base"].reshape(n, 3, 3).transpose(0, 2, 1))


def point_velocity(B, b, X, u):
    """Spatial velocity of the material point with referential position X of body b
    for generalised velocity u = [W (referential angular), v]: v + (R W) x (x_p - x_c)."""
    R, xc, x0 = B["R"][b], B["x"][b], B["x0"][b]
    xp = xc + R @ (X - x0)
    return u[3:] + np.cross(R @ u[:3], xp - xc)


def local_velocity(B, Cc, i, U_of_body):
    m, s = Cc["master"][i], Cc["slave"][i]
    vm = point_velocity(B, m, Cc["mpnt"][i], U_of_body(m))
    vs = point_velocity(B, s, Cc["spnt"][i], U_of_body(s))
    return Cc["base"][i].T @ (vm - vs)


def impulse_response(B, Cc, i, j, k):
    """Local velocity change at contact i from a unit local impulse e_k at contact j."""
    base_j = Cc["base"][j]
    f = base_j[:, k]
    du = {}
    for b, sign, X in ((Cc["master"][j], 1.0, Cc["mpnt"][j]), (Cc["slave"][j], -1.0, Cc["spnt"][j])):
        if B["kind"][b] == 1:
            continue
        R, xc, x0 = B["R"][b], B["x"][b], B["x0"][b]
        xp = xc + R @ (X - x0)
        torque_ref = R.T @ np.cross(xp - xc, sign * f)
        dW = H * B["jinv"][b] @ torque_ref
        dv = H * B["minv"][b] * sign * f
        du[b] = du.get(b, np.zeros(6)) + np.concatenate([dW, dv])
    return local_velocity(B, Cc, i, lambda b: du.get(b, np.zeros(6)) if B["kind"][b] == 0 else np.zeros(6))


@pytest.fixture(scope="module", params=["spheres", "ellipsoids"])
def scene(request):
    if request.param == "spheres":
        s = lattice(64, seed=4)
    else:   # general inertia and contact points off the centre lines
        s = solfec.SOLFEC("DYNAMIC", H, "/tmp/solfec_test")
        solfec.SURFACE_MATERIAL(s, friction=0.3, restitution=0.0)
        solfec.GRAVITY(s, (0.0, 0.0, -9.81))
        solfec.ELLIP_LATTICE(s, 64, 0.4, 0.03, 0.06, 0.005, 2000.0, 4)
        s.step_begin()
    b, c = s.bodies(), s.contacts()
    assert len(c["gap"]) > 20
    O = oracle.assemble(oracle.bodies_from(b), oracle.contacts_from(c), H)
    return b, c, O


def test_free_and_previous_velocities(scene):
    b, c, O = scene
    B, Cc = body_arrays(b), contact_arrays(c)
    n = O["n"]
    for i in range(n):
        Bi = local_velocity(B, Cc, i, lambda k: B["velo"][k])
        Vi = local_velocity(B, Cc, i, lambda k: B["velo0"][k])
        assert np.allclose(O["B"][3 * i:3 * i + 3], Bi, rtol=1e-12, atol=1e-15)
        assert np.allclose(O["V"][3 * i:3 * i + 3], Vi, rtol=1e-12, atol=1e-15)


def test_W_is_the_impulse_response(scene):
    b, c, O = scene
    B, Cc = body_arrays(b), contact_arrays(c)
    n = O["n"]
    Wd = O["Wd"].reshape(n, 3, 3).transpose(0, 2, 1)
    for i in range(n):
        Wi = np.stack([impulse_response(B, Cc, i, i, k) for k in range(3)], axis=1)
        assert np.allclose(Wd[i], Wi, rtol=1e-11, atol=1e-16)
        for q in range(O["rowptr"][i], O["rowptr"][i + 1]):
            j = O["col"][q]
            Wij = np.stack([impulse_response(B, Cc, i, j, k) for k in range(3)], axis=1)
            Wo = O["W"][9 * q:9 * q + 9].reshape(3, 3).T
            assert np.allclose(Wo, Wij, rtol=1e-11, atol=1e-16), (i, j)


def test_adjacency_is_complete_and_excludes_obstacles(scene):
    b, c, O = scene
    n = O["n"]
    m, s, kind = c["master"], c["slave"], b["kind"]
    expect = set()
    for i in range(n):
        for j in range(n):
            if i == j:
                continue
            shared = {m[i], s[i]} & {m[j], s[j]}
            if any(kind[x] == 0 for x in shared):
                expect.add((i, j))
    got = {(i, int(O["col"][q])) for i in range(n) for q in range(O["rowptr"][i], O["rowptr"][i + 1])}
    assert got == expect


def test_symmetry_psd_rho_inverse_energy(scene):
    b, c, O = scene
    n = O["n"]
    dense = np.zeros((3 * n, 3 * n))
    for i in range(n):
        dense[3 * i:3 * i + 3, 3 * i:3 * i + 3] = O["Wd"][9 * i:9 * i + 9].reshape(3, 3).T
        for q in range(O["rowptr"][i], O["rowptr"][i + 1]):
            j = O["col"][q]
            dense[3 * i:3 * i + 3, 3 * j:3 * j + 3] += O["W"][9 * q:9 * q + 9].reshape(3, 3).T
    assert np.allclose(dense, dense.T, atol=1e-14 * np.abs(dense).max())
    assert np.linalg.eigvalsh(0.5 * (dense + dense.T)).min() > -1e-12 * np.abs(dense).max()
    E = 0.0
    for i in range(n):
        Wi = O["Wd"][9 * i:9 * i + 9].reshape(3, 3).T
        assert abs(O["rho"][i] - 1.0 / np.linalg.eigvalsh(Wi).max()) <= 1e-12 * O["rho"][i]
        Ai = O["A"][9 * i:9 * i + 9].reshape(3, 3).T
        assert np.allclose(Ai, np.linalg.inv(Wi), rtol=1e-10)
        Bi = O["B"][3 * i:3 * i + 3]
        E += 0.5 * Bi @ np.linalg.inv(Wi) @ Bi
    assert abs(O["free_energy"] - E) <= 1e-10 * abs(E)


def _two_body_scene(x1, r1, x2=None, r2=None, plane=True):
    s = solfec.SOLFEC("DYNAMIC", H, "/tmp/solfec_test")
    solfec.SURFACE_MATERIAL(s, friction=0.3)
    bulk = solfec.BULK_MATERIAL(s, density=2000.0)
    solfec.BODY(s, "RIGID", solfec.SPHERE(x1, r1, 1, 1), bulk)
    if x2 is not None:
        solfec.BODY(s, "RIGID", solfec.SPHERE(x2, r2, 1, 1), bulk)
    if plane:
        solfec.BODY(s, "OBSTACLE", solfec.PLANE((0, 0, 0), (0, 0, 1), 0), bulk)
    s.step_begin()
    b, c = s.bodies(), s.contacts()
    return b, c, oracle.assemble(oracle.bodies_from(b), oracle.contacts_from(c), H)


def test_sphere_on_plane_closed_form():
    r = 0.05
    b, c, O = _two_body_scene((0.3, 0.2, r - 1e-4), r)
    assert O["n"] == 1
    m = 2000.0 * 4.0 / 3.0 * np.pi * r ** 3
    Wd = O["Wd"].reshape(3, 3).T
    assert np.allclose(Wd, (H / m) * np.diag([3.5, 3.5, 1.0]), rtol=1e-12, atol=1e-18)
    # gap and orientation: normal of the plane, plane is the slave
    assert c["gap"][0] < 0 and c["slave"][0] == 1 and np.allclose(c["base"][6:9], [0, 0, 1])


def test_sphere_pair_closed_form():
    r1, r2 = 0.05, 0.04
    d = r1 + r2 + 5e-7       # inside the detection tolerance, surfaces not overlapping
    b, c, O = _two_body_scene((0.0, 0.0, 0.0), r1, (d / np.sqrt(3), d / np.sqrt(3), d / np.sqrt(3)), r2, plane=False)
    assert O["n"] == 1
    m1 = 2000.0 * 4.0 / 3.0 * np.pi * r1 ** 3
    m2 = 2000.0 * 4.0 / 3.0 * np.pi * r2 ** 3
    inv = 1.0 / m1 + 1.0 / m2
    Wd = O["Wd"].reshape(3, 3).T
    assert np.allclose(Wd, H * np.diag([3.5 * inv, 3.5 * inv, inv]), rtol=1e-9, atol=1e-18)
    # the master is the sphere with the lexicographically larger centre; normal from slave to master
    assert c["master"][0] == 1 and c["slave"][0] == 0
    assert np.allclose(c["base"][6:9], np.ones(3) / np.sqrt(3))
