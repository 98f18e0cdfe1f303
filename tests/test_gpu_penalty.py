"""PENALTY_SOLVER on the GPU (SFGS_Penalty_Solve): every contact's reaction and
velocity against the oracle's spring-dashpot law (pes.c:33-106) applied to the
oracle's W_ii, B and V (ldy.c) from the same carried R and U; the device
loop's material-table lookup against the host loop's per-contact constants,
bit for bit; a sphere settling on a linear spring at the static deflection
m g / k with critical damping.  FP64 throughout: the law is evaluated without
contraction on both sides, so the differences come from the assembled W, B, V
(1e-12 relative) and pow/sqrt rounding; the bar is 1e-10 relative to the
largest reaction / velocity."""
import numpy as np
import pytest

import oracle
import solfec
from scenes import lattice

pytestmark = pytest.mark.gpu

H = 1e-3
RTOL = 1e-10


@pytest.fixture(scope="module")
def dev():
    d = solfec.Device(0)
    yield d
    d.close()


# (implicit, spring, dashpot, hpow): both variants, a linear and a Hertz spring,
# a given dashpot, critical damping (negative dashpot) and none
CASES = [(1, 1e6, 5.0, 1.0), (0, 1e6, 5.0, 1.0), (1, 1e7, -1.0, 1.5), (0, 2e5, 0.0, 1.0), (0, 1e7, -1.0, 1.5)]


@pytest.mark.parametrize("implicit,k,c,p", CASES)
def test_penalty_rows_match_oracle(dev, implicit, k, c, p):
    print(f"[penalty] implicit={implicit} k={k} c={c} p={p}", flush=True)
    s = lattice(343, seed=7)
    b, cc = s.bodies(), s.contacts()
    n, nb = len(cc["master"]), len(b["kind"])
    rng = np.random.default_rng(11)
    # gaps on both sides of zero, on the scale where the spring force k (-g)^p
    # meets the friction cone of the tangential reaction of these velocities
    d = (1e-5 / (k / 1e6)) ** (1.0 / p)
    cc["gap"] = rng.uniform(-2.0 * d, d, n)
    b["velo"] = b["velo"] + rng.normal(scale=0.01, size=6 * nb)
    b["velo0"] = b["velo0"] + rng.normal(scale=0.01, size=6 * nb)
    R0 = rng.normal(scale=5.0, size=(n, 3))
    U0 = rng.normal(size=(n, 3))
    spring = k * (1.0 + 0.5 * rng.random(n))
    dashpot = np.full(n, c)
    hpow = np.full(n, p)
    dev.bodies(b["kind"], b["rot"], b["x0"], b["jinv"], b["minv"], b["velo"], b["velo0"])
    dev.contacts(cc["master"], cc["slave"], cc["mpnt"], cc["spnt"], cc["base"], cc["gap"], cc["friction"],
                 cc["restitution"], cc["cohesion"], R0.ravel(), U0.ravel())
    dev.penalty(H, implicit, spring, dashpot, hpow)
    R, U = dev.get_RU()
    O = oracle.assemble(oracle.bodies_from(b), oracle.contacts_from(cc), H)
    Ro, Uo = np.zeros((n, 3)), np.zeros((n, 3))
    for i in range(n):
        rc, Ro[i], Uo[i] = oracle.penalty_contact(implicit, H, cc["gap"][i], spring[i], dashpot[i], hpow[i],
                                                  cc["friction"][i], O["Wd"][9 * i:9 * i + 9], O["B"][3 * i:3 * i + 3],
                                                  O["V"][3 * i:3 * i + 3], R0[i], U0[i])
        assert rc == 0
    sep = Ro[:, 2] == 0.0
    lim = cc["friction"] * Ro[:, 2]
    slide = ~sep & (np.hypot(Ro[:, 0], Ro[:, 1]) >= lim * (1 - 1e-12))
    print(f"[penalty] {n} contacts: {sep.sum()} separated, {slide.sum()} sliding, {(~sep & ~slide).sum()} sticking;"
          f" max |dR| {np.abs(R - Ro).max():.3e}, max |dU| {np.abs(U - Uo).max():.3e},"
          f" bit-equal R {np.mean(R == Ro):.3f}", flush=True)
    assert sep.sum() > 0 and slide.sum() > 0 and (~sep & ~slide).sum() > 0
    np.testing.assert_array_equal(R[:, 2] == 0.0, sep)
    assert np.abs(R - Ro).max() <= RTOL * np.abs(Ro).max()
    assert np.abs(U - Uo).max() <= RTOL * np.abs(Uo).max()


def spring_scene(seed, implicit):
    """Sphere pairs on the default material, sphere-plane pairs on their own;
    the explicit variant is stable for h sqrt(k / m) well below 1 only."""
    s = solfec.SOLFEC("DYNAMIC", H, "/tmp/solfec_test")
    if implicit:
        solfec.SURFACE_MATERIAL(s, model="SPRING_DASHPOT", friction=0.3, spring=1e6, dashpot=-1.0)
        solfec.SURFACE_MATERIAL(s, 0, 1, model="SPRING_DASHPOT", friction=0.5, spring=3e6, dashpot=20.0, hpow=1.5)
    else:
        solfec.SURFACE_MATERIAL(s, model="SPRING_DASHPOT", friction=0.3, spring=1e4, dashpot=1.0)
        solfec.SURFACE_MATERIAL(s, 0, 1, model="SPRING_DASHPOT", friction=0.5, spring=2e4, dashpot=0.5, hpow=1.5)
    solfec.GRAVITY(s, (0.0, 0.0, -9.81))
    solfec.LATTICE(s, 512, 0.8, 0.045, 0.055, 0.005, 2000.0, seed)
    return s


@pytest.mark.parametrize("variant", ["IMPLICIT", "EXPLICIT"])
def test_penalty_device_loop_equals_host_loop(variant):
    """The host loop passes each contact's constants; the device loop looks
    them up in its material table by surface pair: same bits."""
    print(f"[penalty] device loop against host loop, {variant}", flush=True)
    ps = solfec.PENALTY_SOLVER(variant)
    host, dvc = spring_scene(3, variant == "IMPLICIT"), spring_scene(3, variant == "IMPLICIT")
    dvc.device_loop(True)
    solfec.RUN(host, ps, 0.03)
    solfec.RUN(dvc, ps, 0.03)
    bh, bd = host.bodies(), dvc.bodies()
    for k in ("x", "rot", "velo"):
        assert np.array_equal(bh[k], bd[k]), k
    ch, cd = host.contacts(), dvc.contacts()
    assert len(ch["master"]) == len(cd["master"]) > 100
    for k in ("master", "slave", "gap", "R", "U"):
        assert np.array_equal(ch[k], cd[k]), k
    assert np.abs(ch["R"]).max() > 0


@pytest.mark.parametrize("loop", ["host", "device"])
def test_sphere_settles_on_a_spring(loop):
    """Critical damping: after 0.2 s the sphere rests at the static deflection
    -gap = m g / k and the contact carries its weight."""
    s = solfec.SOLFEC("DYNAMIC", H, "/tmp/solfec_test")
    k = 1e5
    solfec.SURFACE_MATERIAL(s, model="SPRING_DASHPOT", friction=0.5, spring=k, dashpot=-1.0)
    solfec.GRAVITY(s, (0.0, 0.0, -9.81))
    mat = solfec.BULK_MATERIAL(s, density=2000.0)
    r = 0.05
    solfec.BODY(s, "OBSTACLE", solfec.PLANE((0, 0, 0), (0, 0, 1), 0), mat)
    solfec.BODY(s, "RIGID", solfec.SPHERE((0.3, 0.2, r), r, 1, 1), mat)
    s.device_loop(loop == "device")
    solfec.RUN(s, solfec.PENALTY_SOLVER("IMPLICIT"), 0.2)
    b, c = s.bodies(), s.contacts()
    m = 2000.0 * 4.0 / 3.0 * np.pi * r ** 3
    assert len(c["master"]) == 1
    np.testing.assert_allclose(-c["gap"][0], m * 9.81 / k, rtol=1e-8)
    np.testing.assert_allclose(b["x"][3:6], [0.3, 0.2, r - m * 9.81 / k], rtol=0, atol=1e-11)
    np.testing.assert_allclose(b["velo"][6:12], 0.0, atol=1e-10)
    np.testing.assert_allclose(c["R"], [0.0, 0.0, m * 9.81], rtol=1e-8, atol=1e-12)


def test_penalty_store_refuses_the_gauss_seidel_calls(dev):
    s = lattice(27, seed=2)
    b, cc = s.bodies(), s.contacts()
    n = len(cc["master"])
    dev.bodies(b["kind"], b["rot"], b["x0"], b["jinv"], b["minv"], b["velo"], b["velo0"])
    dev.contacts(cc["master"], cc["slave"], cc["mpnt"], cc["spnt"], cc["base"], cc["gap"], cc["friction"],
                 cc["restitution"], cc["cohesion"], cc["R"], cc["U"])
    dev.penalty(H, 1, np.full(n, 1e6), np.zeros(n), np.ones(n))
    with pytest.raises(solfec.SfgsError):
        dev.merit()
    with pytest.raises(solfec.SfgsError):
        dev.locdyn()
    with pytest.raises(solfec.SfgsError):       # no material table on this context
        dev.penalty(H, 1)
    dev.assemble(H)                             # a full assembly makes the store a system again
    dev.merit()


@pytest.mark.parametrize("solver", ["penalty", "newton"])
def test_restart_with_other_solvers_is_bit_exact(tmp_path, solver):
    """Five steps, a state file, five more; a fresh domain of the same scene that
    reads the file and runs the same five steps ends in the same bits (device loop)."""
    def make():
        s = spring_scene(4, True)
        s.device_loop(True)
        return s
    sv = solfec.PENALTY_SOLVER() if solver == "penalty" else solfec.NEWTON_SOLVER(meritval=1e-8, maxiter=100)
    a = make()
    solfec.RUN(a, sv, 0.005)
    a.write_state(tmp_path / "s.xdr")
    solfec.RUN(a, sv, 0.005)
    b = make()
    b.read_state(tmp_path / "s.xdr")
    solfec.RUN(b, sv, 0.005)
    ba, bb = a.bodies(), b.bodies()
    for k in ("x", "rot", "velo"):
        assert np.array_equal(ba[k], bb[k]), k
    ca, cb = a.contacts(), b.contacts()
    assert len(ca["master"]) == len(cb["master"]) > 100
    for k in ("R", "U", "gap"):
        assert np.array_equal(ca[k], cb[k]), k
