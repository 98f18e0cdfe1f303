"""The device loop (SOLFEC_Device_Loop: DOM_Update_Begin / DOM_Update_End on
the GPU, bodies and contacts kept in HBM) against the host loop, which runs the
same geometry and integration source on the CPU with the same GPU solve: the
two must agree bit for bit, step after step, for spheres and for ellipsoids,
through host edits and switches between the loops.  Also the device's contact
detection against the oracle's brute force."""
import numpy as np
import pytest

import oracle
import solfec

pytestmark = pytest.mark.gpu

BODY_KEYS = ("x", "rot", "velo", "velo0")
CON_KEYS = ("master", "slave", "mpnt", "spnt", "base", "gap", "friction", "R", "U")


def scene(n, L, seed, shape="sphere", friction=0.3):
    s = solfec.SOLFEC("DYNAMIC", 1e-3, "/tmp/solfec_test")
    solfec.SURFACE_MATERIAL(s, friction=friction, restitution=0.0)
    solfec.GRAVITY(s, (0.0, 0.0, -9.81))
    if shape == "ellipsoid":
        solfec.ELLIP_LATTICE(s, n, L, 0.03, 0.06, 0.005, 2000.0, seed)
    else:
        solfec.LATTICE(s, n, L, 0.045, 0.055, 0.005, 2000.0, seed)
    return s


def step(s, gs):
    assert s.step_begin() == 0, s.lib.SOLFEC_Last_Error(s.sol).decode()
    assert s.step_solve(gs) == 0, s.lib.SOLFEC_Last_Error(s.sol).decode()
    assert s.step_end() == 0, s.lib.SOLFEC_Last_Error(s.sol).decode()


def same_state(a, b, what):
    ba, bb = a.bodies(), b.bodies()
    for k in BODY_KEYS:
        d = np.abs(ba[k] - bb[k]).max() if len(ba[k]) else 0.0
        assert np.array_equal(ba[k], bb[k]), f"{what}: bodies {k} differ by {d:.3e}"
    ca, cb = a.contacts(), b.contacts()
    assert len(ca["master"]) == len(cb["master"]), f"{what}: {len(ca['master'])} against {len(cb['master'])} contacts"
    for k in CON_KEYS:
        d = np.abs(ca[k] - cb[k]).max() if len(ca[k]) else 0.0
        assert np.array_equal(ca[k], cb[k]), f"{what}: contacts {k} differ by {d:.3e}"
    return len(ca["master"])


@pytest.mark.parametrize("shape,n,L,steps", [("sphere", 1000, 1.0, 60), ("ellipsoid", 343, 0.7, 40)])
def test_device_loop_equals_host_loop(shape, n, L, steps):
    print(f"[dom] {shape} n={n} steps={steps}", flush=True)
    host, dev = scene(n, L, 4, shape), scene(n, L, 4, shape)
    dev.device_loop(True)
    gs = solfec.GAUSS_SEIDEL_SOLVER(1e-4, 1000)
    for k in range(steps):
        step(host, gs)
        step(dev, gs)
        if (k + 1) % 20 == 0:
            nc = same_state(host, dev, f"step {k + 1}")
            print(f"[dom] step {k + 1}: {nc} contacts, identical", flush=True)
    assert host.ncon > n // 2


def test_device_loop_edits_and_switches():
    host, dev = scene(512, 0.8, 6), scene(512, 0.8, 6)
    dev.device_loop(True)
    gs = solfec.GAUSS_SEIDEL_SOLVER(1e-4, 1000)
    for k in range(10):
        step(host, gs)
        step(dev, gs)
    for s in (host, dev):                 # an edit on the host between steps
        s.set_velocity(17, [0.0, 5.0, 0.0, 1.0, -2.0, 0.5])
    for k in range(10):
        step(host, gs)
        step(dev, gs)
    same_state(host, dev, "after an edit")
    dev.device_loop(False)                # back to the host loop, then to the device again
    for k in range(5):
        step(host, gs)
        step(dev, gs)
    dev.device_loop(True)
    solfec.RUN(host, gs, 0.005)
    solfec.RUN(dev, gs, 0.005)
    same_state(host, dev, "after switching and RUN")


def test_device_detection_matches_brute_force():
    n, L = 343, 0.6                       # overlapping neighbours
    s = scene(n, L, 5)
    s.device_loop(True)
    assert s.step_begin() == 0
    b, c = s.bodies(), s.contacts()
    shape = np.r_[np.zeros(n, np.int32), np.ones(6, np.int32)]
    plane = np.zeros((n + 6, 6))
    plane[n:] = [[0, 0, 0, 1, 0, 0], [L, 0, 0, -1, 0, 0], [0, 0, 0, 0, 1, 0], [0, L, 0, 0, -1, 0],
                 [0, 0, 0, 0, 0, 1], [0, 0, L, 0, 0, -1]]
    O = oracle.detect(shape, b["radius"], plane.ravel(), b["x"])
    print(f"[dom] {len(O['master'])} contacts by brute force, {len(c['master'])} on the device", flush=True)
    np.testing.assert_array_equal(c["master"], O["master"])
    np.testing.assert_array_equal(c["slave"], O["slave"])
    np.testing.assert_allclose(c["base"].reshape(-1, 3, 3)[:, 2, :], O["normal"], rtol=0, atol=1e-14)
    np.testing.assert_allclose(c["gap"], O["gap"], rtol=0, atol=1e-15)
    assert np.all(c["R"] == 0) and np.all(c["U"] == 0)


def test_device_loop_dump_equals_host(tmp_path):
    host, dev = scene(216, 0.6, 8), scene(216, 0.6, 8)
    dev.device_loop(True)
    gs = solfec.GAUSS_SEIDEL_SOLVER(1e-4, 1000)
    for k in range(3):
        step(host, gs)
        step(dev, gs)
    solfec.LOCDYN_DUMP(host, tmp_path / "h.txt")
    solfec.LOCDYN_DUMP(dev, tmp_path / "d.txt")
    assert (tmp_path / "h.txt").read_text() == (tmp_path / "d.txt").read_text()


@pytest.mark.parametrize("loop", ["host", "device"])
def test_restart_from_state_file_is_bit_exact(tmp_path, loop):
    """Ten steps, a state file, ten more steps; a fresh domain of the same scene
    that reads the file and runs the same ten steps ends in the same bits."""
    gs = solfec.GAUSS_SEIDEL_SOLVER(1e-4, 1000)
    a = scene(512, 0.8, 9)
    a.device_loop(loop == "device")
    for k in range(10):
        step(a, gs)
    a.write_state(tmp_path / "s.xdr")
    for k in range(10):
        step(a, gs)
    b = scene(512, 0.8, 9)
    b.device_loop(loop == "device")
    b.read_state(tmp_path / "s.xdr")
    for k in range(10):
        step(b, gs)
    nc = same_state(a, b, f"restart ({loop} loop)")
    assert a.time == b.time
    print(f"[dom] restart, {loop} loop: {nc} contacts identical after 10 steps", flush=True)


@pytest.mark.parametrize("loop", ["host", "device"])
def test_sphere_resting_on_a_plane(loop):
    """A sphere at rest on a plane stays at rest; the contact carries its weight:
    R = (0, 0, m g) in the contact frame (R is a force, u += h M^-1 H' R)."""
    s = solfec.SOLFEC("DYNAMIC", 1e-3, "/tmp/solfec_test")
    solfec.SURFACE_MATERIAL(s, friction=0.5, restitution=0.0)
    solfec.GRAVITY(s, (0.0, 0.0, -9.81))
    mat = solfec.BULK_MATERIAL(s, density=2000.0)
    r = 0.05
    solfec.BODY(s, "OBSTACLE", solfec.PLANE((0, 0, 0), (0, 0, 1), 0), mat)
    solfec.BODY(s, "RIGID", solfec.SPHERE((0.3, 0.2, r), r, 1, 1), mat)
    s.device_loop(loop == "device")
    gs = solfec.GAUSS_SEIDEL_SOLVER(1e-10, 100)
    solfec.RUN(s, gs, 0.05)
    b, c = s.bodies(), s.contacts()
    m = 2000.0 * 4.0 / 3.0 * np.pi * r ** 3
    assert len(c["master"]) == 1
    np.testing.assert_allclose(b["x"][3:6], [0.3, 0.2, r], atol=1e-9)
    np.testing.assert_allclose(b["velo"][6:12], 0.0, atol=1e-9)
    np.testing.assert_allclose(c["R"], [0.0, 0.0, m * 9.81], rtol=1e-9, atol=1e-12)
    assert s.time == pytest.approx(0.05)


def test_device_loop_without_convex_bodies():
    """Planes only, and then nothing at all: the device loop runs empty steps."""
    s = solfec.SOLFEC("DYNAMIC", 1e-3, "/tmp/solfec_test")
    mat = solfec.BULK_MATERIAL(s, density=1000.0)
    solfec.BODY(s, "OBSTACLE", solfec.PLANE((0, 0, 0), (0, 0, 1), 0), mat)
    s.device_loop(True)
    solfec.RUN(s, solfec.GAUSS_SEIDEL_SOLVER(1e-6, 100), 0.003)
    assert s.ncon == 0 and s.time == pytest.approx(0.003)
    e = solfec.SOLFEC("DYNAMIC", 1e-3, "/tmp/solfec_test")
    e.device_loop(True)
    solfec.RUN(e, solfec.GAUSS_SEIDEL_SOLVER(1e-6, 100), 0.002)
    assert e.nbod == 0 and e.ncon == 0
