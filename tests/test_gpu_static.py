"""QUASI_STATIC analysis (SOLFEC('QUASI_STATIC', ...), dom->dynamic = 0) on the
GPU: the rigid static step (BODY_Static_Step_Begin/End, bod.c:1588-1645) in the
host and the device loop, and the local dynamics with the static gap shift
max(gap, 0)/h (dbs.c semismooth_newton with dynamic = 0, no restitution).

* A sphere resting on a plane: every step starts it from rest with velocity
  h g, the contact answers with R_N = m g exactly, the velocity returns to 0
  and the sphere stays put (closed form).
* The device loop equals the host loop bit for bit over a settling lattice.
* The Gauss-Seidel solve of a static system equals the oracle's
  (OR_SYSTEM with dynamic = 0) in the GPU's colour order."""
import numpy as np
import pytest

import oracle
import solfec
from scenes import lattice

pytestmark = pytest.mark.gpu

H = 1e-3


@pytest.fixture(scope="module")
def dev():
    d = solfec.Device(0)
    yield d
    d.close()


@pytest.mark.parametrize("loop", ["host", "device"])
def test_quasi_static_sphere_rests_on_a_plane(loop):
    print(f"[static] sphere on a plane, {loop} loop", flush=True)
    s = solfec.SOLFEC("QUASI_STATIC", H, "/tmp/solfec_test")
    solfec.SURFACE_MATERIAL(s, friction=0.5)
    solfec.GRAVITY(s, (0.0, 0.0, -9.81))
    mat = solfec.BULK_MATERIAL(s, density=2000.0)
    solfec.BODY(s, "OBSTACLE", solfec.PLANE((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 0), mat)
    r = 0.05
    solfec.BODY(s, "RIGID", solfec.SPHERE((0.3, 0.2, r - 1e-6), r, 1, 1), mat)
    s.device_loop(loop == "device")
    solfec.RUN(s, solfec.GAUSS_SEIDEL_SOLVER(1e-10, 100), 0.02)
    b, c = s.bodies(), s.contacts()
    m = 2000.0 * 4.0 / 3.0 * np.pi * r ** 3
    print(f"[static] t={s.time:.3f}: R {c['R']}, velocity {b['velo'][6:12]}", flush=True)
    assert len(c["master"]) == 1
    np.testing.assert_allclose(c["R"][2], m * 9.81, rtol=1e-9)
    np.testing.assert_allclose(c["R"][:2], 0.0, atol=1e-9)
    assert np.abs(b["velo"][6:12]).max() <= 1e-12
    np.testing.assert_allclose(b["x"][3:6], [0.3, 0.2, r - 1e-6], rtol=0, atol=1e-12)


def test_quasi_static_device_loop_equals_host_loop():
    print("[static] lattice, device loop against host loop", flush=True)

    def make():
        return lattice(512, L=0.8, seed=3, steps_begin=0, analysis="QUASI_STATIC")
    gs = solfec.GAUSS_SEIDEL_SOLVER(1e-6, 1000)
    host, dvc = make(), make()
    dvc.device_loop(True)
    for duration in (0.002, 0.018):
        solfec.RUN(host, gs, duration)
        solfec.RUN(dvc, gs, duration)
        bh, bd = host.bodies(), dvc.bodies()
        for k in ("x", "rot", "velo", "velo0"):
            assert np.array_equal(bh[k], bd[k]), k
        ch, cd = host.contacts(), dvc.contacts()
        print(f"[static] t={host.time:.3f}: {len(ch['master'])} contacts, max |v| "
              f"{np.abs(bh['velo']).max():.3e}", flush=True)
        assert len(ch["master"]) == len(cd["master"]) > 100
        for k in ("master", "slave", "gap", "R", "U", "mpnt", "spnt"):
            assert np.array_equal(ch[k], cd[k]), k
    # the static scheme carries no velocity: the previous velocities stay as they were set
    assert np.all(bh["velo0"] == 0.0)


def test_quasi_static_solve_matches_oracle(dev):
    print("[static] Gauss-Seidel on a static system against the oracle", flush=True)
    s = lattice(1000, seed=2, analysis="QUASI_STATIC")
    b, c = s.bodies(), s.contacts()
    n = len(c["master"])
    # a third of the contacts open by up to 10 um: the static shift max(gap, 0)/h acts on them
    rng = np.random.default_rng(2)
    c = dict(c)
    c["gap"] = np.where(rng.random(n) < 0.3, 1e-5 * rng.random(n), c["gap"])
    dev.bodies(b["kind"], b["rot"], b["x0"], b["jinv"], b["minv"], b["velo"], b["velo0"])
    dev.contacts(c["master"], c["slave"], c["mpnt"], c["spnt"], c["base"], c["gap"], c["friction"],
                 c["restitution"], c["cohesion"], np.zeros(3 * n), np.zeros(3 * n))
    dev.assemble(H, dynamic=0)
    colours = dev.locdyn()["colour"]
    kw = dict(epsilon=1e-8, maxiter=3000, meritval=1e-10)
    rep, _, _ = dev.solve(solfec.make_param(**kw))
    Rg, Ug = dev.get_RU()
    O = oracle.assemble(oracle.bodies_from(b), oracle.contacts_from(c), H)
    order, _ = oracle.colour_order(colours, int(colours.max()) + 1)
    rc, orep, Ro, Uo, _, _ = oracle.gs_solve(oracle.System(O, dynamic=0), oracle.gs_param(**kw), order,
                                             np.zeros(3 * n), np.zeros(3 * n), history=False)
    Ro, Uo = np.asarray(Ro).reshape(n, 3), np.asarray(Uo).reshape(n, 3)
    err = np.abs(Rg - Ro).max() / np.abs(Ro).max()
    print(f"[static] {n} contacts, {int((O['gap'] > 0).sum())} with gap > 0: iters {rep.iters} / {orep.iters}, "
          f"merit {rep.merit:.3e} / {orep.merit:.3e}, max |dR|/max|R| {err:.2e}", flush=True)
    assert abs(rep.iters - orep.iters) <= 1
    assert err < 1e-6
    assert np.abs(Ug - Uo).max() <= 1e-6 * np.abs(Uo).max()
