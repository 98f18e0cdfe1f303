"""Cohesive contacts on the GPU: the Gauss-Seidel solve with cohesion states
against the oracle's colour-ordered solve on the same system (the cone shifted
by the cohesion, scf.c:136-174), update_cohesion (ldy.c:69-95) on the device
against the oracle's on the same reactions (flags, bit for bit), the cohesive
spring-dashpot law (pes.c:59-99) against the oracle per contact, and the
physics of a sphere glued under a ceiling: it hangs while its weight is below
the cohesive force and falls, with the cohesion lost, when it is above, in the
host and the device loop alike, and the two loops agree bit for bit."""
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


def _cohesive_contacts(c, seed):
    """Half the contacts cohesive, with states log-uniform over [1e-3, 10] N."""
    rng = np.random.default_rng(seed)
    n = len(c["master"])
    coh = np.where(rng.random(n) < 0.5, 10.0 ** rng.uniform(-3.0, 1.0, n), 0.0)
    c = dict(c)
    c["cohesion"] = coh
    return c


@pytest.mark.parametrize("n,seed", [(343, 5), (1000, 1)])
def test_gs_with_cohesion_and_update_cohesion_match_oracle(dev, n, seed):
    print(f"[cohesion] gs n={n} seed={seed}", flush=True)
    s = lattice(n, seed=seed)
    b, c = s.bodies(), s.contacts()
    c = _cohesive_contacts(c, seed)
    nc = len(c["master"])
    dev.bodies(b["kind"], b["rot"], b["x0"], b["jinv"], b["minv"], b["velo"], b["velo0"])
    dev.contacts(c["master"], c["slave"], c["mpnt"], c["spnt"], c["base"], c["gap"], c["friction"],
                 c["restitution"], c["cohesion"], c["R"], c["U"])
    dev.assemble(H)
    kw = dict(epsilon=1e-6, maxiter=3000, meritval=1e-6)
    rep, _, _ = dev.solve(solfec.make_param(**kw))
    Rg, Ug = dev.get_RU()
    colours = dev.locdyn()["colour"]
    order, _ = oracle.colour_order(colours, int(colours.max()) + 1)
    O = oracle.assemble(oracle.bodies_from(b), oracle.contacts_from(c), H)
    rc, orep, Ro, Uo, _, _ = oracle.gs_solve(oracle.System(O), oracle.gs_param(**kw), order, np.zeros(3 * nc),
                                             np.zeros(3 * nc))
    Ro = np.asarray(Ro).reshape(nc, 3)
    scale = np.abs(Ro).max()
    err = np.abs(Rg - Ro).max() / scale
    print(f"[cohesion] {nc} contacts, iters {rep.iters}/{orep.iters}, max |dR|/max|R| {err:.2e},"
          f" tension rows {(Ro[:, 2] < 0).sum()}", flush=True)
    assert (Ro[:, 2] < 0).sum() > 0            # cohesion carries tension somewhere
    assert err < 1e-6
    dev.update_cohesion()
    cg = dev.cohesion()
    co = oracle.update_cohesion(Rg, c["friction"], c["cohesion"])
    kept, broke = (cg > 0).sum(), ((c["cohesion"] > 0) & (cg == 0)).sum()
    print(f"[cohesion] states: {kept} kept, {broke} broken", flush=True)
    np.testing.assert_array_equal(cg, co)
    assert kept > 0 and broke > 0


def test_penalty_with_cohesion_matches_oracle(dev):
    print("[cohesion] penalty rows", flush=True)
    s = lattice(343, seed=7)
    b, c = s.bodies(), s.contacts()
    n, nb = len(c["master"]), len(b["kind"])
    rng = np.random.default_rng(13)
    c = _cohesive_contacts(c, 13)
    c["gap"] = rng.uniform(-2e-5, 1e-5, n)
    # velocities of 1 mm/s against a stiff dashpot: normal reactions of a few N
    # either way, tangential ones well inside the cone (W_TT ~ 7e-3 s/kg), so
    # that tension is held, or breaks the cohesion, on either side of the states
    b["velo"] = b["velo"] + rng.normal(scale=0.001, size=6 * nb)
    b["velo0"] = b["velo0"] + rng.normal(scale=0.001, size=6 * nb)
    R0 = rng.normal(scale=0.5, size=(n, 3))
    U0 = rng.normal(size=(n, 3))
    spring, dashpot, hpow = np.full(n, 1e6), np.full(n, 5000.0), np.ones(n)
    dev.bodies(b["kind"], b["rot"], b["x0"], b["jinv"], b["minv"], b["velo"], b["velo0"])
    dev.contacts(c["master"], c["slave"], c["mpnt"], c["spnt"], c["base"], c["gap"], c["friction"],
                 c["restitution"], c["cohesion"], R0.ravel(), U0.ravel())
    dev.penalty(H, 1, spring, dashpot, hpow)
    R, U = dev.get_RU()
    cg = dev.cohesion()
    O = oracle.assemble(oracle.bodies_from(b), oracle.contacts_from(c), H)
    Ro, Uo, co = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
    for i in range(n):
        rc, Ro[i], Uo[i], co[i] = oracle.penalty_contact(1, H, c["gap"][i], spring[i], dashpot[i], hpow[i],
                                                         c["friction"][i], O["Wd"][9 * i:9 * i + 9],
                                                         O["B"][3 * i:3 * i + 3], O["V"][3 * i:3 * i + 3], R0[i],
                                                         U0[i], cohesion=c["cohesion"][i])
        assert rc == 0
    held = (c["cohesion"] > 0) & (Ro[:, 2] < 0) & (co > 0)
    modeI = (c["cohesion"] > 0) & (co == 0) & (Ro[:, 2] == -c["cohesion"])
    print(f"[cohesion] {n} contacts: {held.sum()} in tension and cohesive, {modeI.sum()} broken in mode I,"
          f" {((c['cohesion'] > 0) & (co == 0)).sum()} broken in all; max |dR| {np.abs(R - Ro).max():.3e}", flush=True)
    assert held.sum() > 0 and modeI.sum() > 0
    np.testing.assert_array_equal(cg, co)
    assert np.abs(R - Ro).max() <= 1e-10 * np.abs(Ro).max()
    assert np.abs(U - Uo).max() <= 1e-10 * np.abs(Uo).max()


def hanging(cohesion, loop, solver):
    s = solfec.SOLFEC("DYNAMIC", H, "/tmp/solfec_test")
    if solver == "penalty":
        solfec.SURFACE_MATERIAL(s, model="SPRING_DASHPOT", friction=0.5, cohesion=cohesion, spring=1e6,
                                dashpot=-1.0)
    else:
        solfec.SURFACE_MATERIAL(s, friction=0.5, cohesion=cohesion, restitution=0.0)
    solfec.GRAVITY(s, (0.0, 0.0, -9.81))
    mat = solfec.BULK_MATERIAL(s, density=2000.0)
    r = 0.05
    solfec.BODY(s, "OBSTACLE", solfec.PLANE((0, 0, 1.0), (0, 0, -1), 0), mat)
    solfec.BODY(s, "RIGID", solfec.SPHERE((0.3, 0.2, 1.0 - r), r, 1, 1), mat)
    s.device_loop(loop == "device")
    sv = solfec.PENALTY_SOLVER("IMPLICIT") if solver == "penalty" else solfec.GAUSS_SEIDEL_SOLVER(1e-10, 100)
    solfec.RUN(s, sv, 0.05)
    return s


@pytest.mark.parametrize("solver", ["gs", "penalty"])
@pytest.mark.parametrize("loop", ["host", "device"])
def test_sphere_glued_under_a_ceiling(loop, solver):
    """Weight m g = 10.27 N: a cohesive force of 20 N holds the sphere (the
    contact keeps R_N = -m g and its cohesion), 5 N does not.  With the penalty
    law the tension reaches -5 N, the contact loses its cohesion (mode I,
    pes.c:76-82) and is removed once the sphere has left.  With Gauss-Seidel the
    semismooth Newton step of an open contact drives R to 0, not to -cohesion
    (dbs.c:265-279): the sphere falls freely, R = 0 leaves the state cohesive
    (neither decohesion mode of ldy.c:85-87 applies) and the separated contact
    is kept (dom.c:452)."""
    print(f"[cohesion] hanging sphere, {loop} loop, {solver}", flush=True)
    r = 0.05
    m = 2000.0 * 4.0 / 3.0 * np.pi * r ** 3
    s = hanging(20.0, loop, solver)
    b, c = s.bodies(), s.contacts()
    assert len(c["master"]) == 1 and c["cohesion"][0] == 20.0
    if solver == "gs":
        np.testing.assert_allclose(b["x"][3:6], [0.3, 0.2, 1.0 - r], atol=1e-9)
        np.testing.assert_allclose(c["R"][2], -m * 9.81, rtol=1e-6)
    else:
        # the spring pulls nothing at a non-negative gap: the critical dashpot
        # (2 sqrt(h k m) = 2047 N s/m) lets the sphere creep at m g / c = 5 mm/s
        assert abs(b["velo"][11]) < 0.01 and abs(b["x"][5] - (1.0 - r)) < 1e-3
        assert -20.0 < c["R"][2] < 0.0
    s = hanging(5.0, loop, solver)
    b, c = s.bodies(), s.contacts()
    assert b["x"][5] < 1.0 - r - 0.5 * 9.81 * 0.04 ** 2
    if solver == "gs":
        assert len(c["master"]) == 1 and c["cohesion"][0] == 5.0
        np.testing.assert_allclose(c["R"], 0.0, atol=1e-12)
        np.testing.assert_allclose(b["velo"][11], -9.81 * 0.05, rtol=1e-12)
    else:
        assert len(c["master"]) == 0
        assert b["velo"][11] < -9.81 * 0.04


@pytest.mark.parametrize("solver", ["gs", "penalty"])
def test_cohesive_device_loop_equals_host_loop(solver):
    """A settling lattice whose contacts created at time 0 are cohesive: the
    contacts that separate stay in the list while cohesive, both loops in the
    same bits, cohesion states included."""
    print(f"[cohesion] device loop against host loop, {solver}", flush=True)

    def make():
        s = solfec.SOLFEC("DYNAMIC", H, "/tmp/solfec_test")
        if solver == "penalty":
            solfec.SURFACE_MATERIAL(s, model="SPRING_DASHPOT", friction=0.3, cohesion=0.05, spring=1e6,
                                    dashpot=-1.0)
        else:
            solfec.SURFACE_MATERIAL(s, friction=0.3, cohesion=0.05, restitution=0.0)
        solfec.GRAVITY(s, (0.0, 0.0, -9.81))
        solfec.LATTICE(s, 512, 0.8, 0.045, 0.055, 0.005, 2000.0, 3)
        return s
    sv = solfec.PENALTY_SOLVER("IMPLICIT") if solver == "penalty" else solfec.GAUSS_SEIDEL_SOLVER(1e-6, 1000)
    host, dvc = make(), make()
    dvc.device_loop(True)
    cohesive = []
    for duration in (0.002, 0.028):           # two steps, then on to 30
        solfec.RUN(host, sv, duration)
        solfec.RUN(dvc, sv, duration)
        bh, bd = host.bodies(), dvc.bodies()
        for k in ("x", "rot", "velo"):
            assert np.array_equal(bh[k], bd[k]), k
        ch, cd = host.contacts(), dvc.contacts()
        cohesive.append(int((ch["cohesion"] > 0).sum()))
        print(f"[cohesion] t={host.time:.3f}: {len(ch['master'])} contacts, {cohesive[-1]} cohesive", flush=True)
        assert len(ch["master"]) == len(cd["master"]) > 100
        for k in ("master", "slave", "gap", "R", "U", "cohesion", "mpnt", "spnt"):
            assert np.array_equal(ch[k], cd[k]), k
    assert cohesive[0] > 0


@pytest.mark.parametrize("loop", ["host", "device"])
def test_newton_keeps_spring_dashpot_cohesion(loop):
    """update_cohesion skips SPRING_DASHPOT contacts (ldy.c:78), which a
    NEWTON_SOLVER step solves as SIGNORINI_COULOMB ones (nts.c:1072-1081).
    Spheres sliding on a floor (1 m/s, friction 0.3), glued at time 0 with a
    cohesion of 0.5 N: surface 1 meets the floor through a SIGNORINI_COULOMB
    material, surface 2 through a SPRING_DASHPOT one.  After one step the
    SIGNORINI_COULOMB contacts hold the states that update_cohesion gives for
    their reactions (the oracle's, bit for bit; the sliding ones break in mode
    II), and the SPRING_DASHPOT contacts keep 0.5 N although the same test
    would break some of them."""
    print(f"[cohesion] NEWTON, SPRING_DASHPOT contacts keep their cohesion, {loop} loop", flush=True)
    s = solfec.SOLFEC("DYNAMIC", H, "/tmp/solfec_test")
    solfec.SURFACE_MATERIAL(s, 0, 1, friction=0.3, cohesion=0.5)
    solfec.SURFACE_MATERIAL(s, 0, 2, model="SPRING_DASHPOT", friction=0.3, cohesion=0.5, spring=1e6, dashpot=-1.0)
    solfec.GRAVITY(s, (0.0, 0.0, -9.81))
    mat = solfec.BULK_MATERIAL(s, density=2000.0)
    solfec.BODY(s, "OBSTACLE", solfec.PLANE((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 0), mat)
    r = 0.05
    surf = []
    for i in range(10):
        for j in range(10):
            sid = 1 + (i + j) % 2
            body = solfec.BODY(s, "RIGID", solfec.SPHERE((0.12 * i, 0.12 * j, r - 1e-4), r, 1, sid), mat)
            solfec.INITIAL_VELOCITY(body, (1.0, 0.0, 0.0), (0.0, 0.0, 0.0))
            surf.append(sid)
    s.device_loop(loop == "device")
    solfec.RUN(s, solfec.NEWTON_SOLVER(meritval=1e-12, maxiter=500), H)
    c = s.contacts()
    assert len(c["master"]) == 100
    sphere = np.where(c["master"] > 0, c["master"], c["slave"]) - 1
    sd = np.asarray(surf)[sphere] == 2
    rule = oracle.update_cohesion(c["R"], c["friction"], np.full(100, 0.5))
    print(f"[cohesion] SIGNORINI_COULOMB: {(c['cohesion'][~sd] == 0).sum()} of {(~sd).sum()} broken; "
          f"SPRING_DASHPOT: {(rule[sd] == 0).sum()} of {sd.sum()} would break, "
          f"{(c['cohesion'][sd] == 0.5).sum()} kept", flush=True)
    assert sd.sum() == 50
    np.testing.assert_array_equal(c["cohesion"][~sd], rule[~sd])
    assert (c["cohesion"][~sd] == 0).sum() > 0
    assert np.all(c["cohesion"][sd] == 0.5)
    assert (rule[sd] == 0).sum() > 0
