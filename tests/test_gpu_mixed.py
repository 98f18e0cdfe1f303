"""Mixed precision (FP32 W/B/R sweeps with FP32 accumulation; FP64 merit,
errors, termination and body update; then FP64 sweeps under the reference's
termination test) against the FP64 oracle on the same system, same colour
order, same R0.

Packings are hyperstatic, so two converged R of one system need not agree
row by row.  What the contact problem fixes is checked instead: the relative
velocities U = B + W R and the velocity of every body after the step (the
per-body impulse sum of H' R, bod.c:418-435, 1446-1447), both within 1e-6 of
their largest magnitude, and the FP64 merit of the returned state at or below
meritval.  R itself is checked within 1e-6 of max|R| where it is unique: one
layer of spheres on a plane, each carried by its own contact."""
import numpy as np
import pytest

import oracle
import solfec
from scenes import lattice

pytestmark = pytest.mark.gpu

TOL = 1e-6


@pytest.fixture(scope="module")
def dev():
    d = solfec.Device(0)
    yield d
    d.close()


def _both(dev, s, kw):
    """GPU MIXED solve from the bodies and the oracle's FP64 solve of the same
    W, B in the GPU's colour order, both from R = U = 0; returns R, U and the
    body velocities after the step for both, and the GPU report."""
    b, c = s.bodies(), s.contacts()
    n = len(c["master"])
    z = np.zeros(3 * n)
    dev.bodies(b["kind"], b["rot"], b["x0"], b["jinv"], b["minv"], b["velo"], b["velo0"])
    dev.contacts(c["master"], c["slave"], c["mpnt"], c["spnt"], c["base"], c["gap"], c["friction"],
                 c["restitution"], c["cohesion"], z, z)
    dev.assemble(1e-3)
    colours = dev.locdyn()["colour"]
    rep, rer, mer = dev.solve(solfec.make_param(**kw, precision=solfec.device.MIXED, refine=2))
    Rg, Ug = dev.get_RU()
    vg, _ = dev.body_update(1e-3, len(b["kind"]))
    O = oracle.assemble(oracle.bodies_from(b), oracle.contacts_from(c), 1e-3)
    order, _ = oracle.colour_order(colours, int(colours.max()) + 1)
    _, orep, Ro, Uo, _, _ = oracle.gs_solve(oracle.System(O), oracle.gs_param(**kw), order, z, z, history=False)
    vo, _ = oracle.body_update(oracle.bodies_from(b), oracle.contacts_from(c), Ro.ravel(), 1e-3)
    m64 = oracle.merit(oracle.System(O), Rg.ravel(), Ug.ravel(), 1)[0]
    return rep, rer, mer, orep, Rg, Ro, Ug, Uo, vg, vo, m64


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("shape,n,seed", [("sphere", 343, 3), ("sphere", 1000, 1), ("ellipsoid", 343, 4)])
def test_mixed_matches_fp64_on_what_the_problem_fixes(dev, shape, n, seed):
    print(f"[mixed] {shape} n={n} seed={seed}", flush=True)
    if shape == "ellipsoid":
        s = solfec.SOLFEC("DYNAMIC", 1e-3, "/tmp/solfec_test")
        solfec.SURFACE_MATERIAL(s, friction=0.3)
        solfec.GRAVITY(s, (0.0, 0.0, -9.81))
        solfec.ELLIP_LATTICE(s, n, 0.6, 0.03, 0.06, 0.005, 2000.0, seed)
        s.step_begin()
    else:
        s = lattice(n, seed=seed)
    kw = dict(epsilon=1e-9, maxiter=20000, meritval=1e-12)
    rep, rer, mer, orep, Rg, Ro, Ug, Uo, vg, vo, m64 = _both(dev, s, kw)
    eU, ev, eR = _rel(Ug, Uo), _rel(vg, vo), _rel(Rg, Ro)
    print(f"[mixed] iters fp64={orep.iters} mixed={rep.iters} merit={rep.merit:.3e} (oracle {orep.merit:.3e}) "
          f"U {eU:.2e} velocities {ev:.2e} R {eR:.2e}", flush=True)
    assert orep.iters < kw["maxiter"] and rep.iters < kw["maxiter"]
    assert rep.merit <= kw["meritval"] and rep.rel_error <= kw["epsilon"]
    assert abs(m64 - rep.merit) <= 1e-6 * max(m64, 1e-300) + 1e-18   # the reported merit is the FP64 merit
    assert len(rer) == rep.iters == len(mer) and np.all(np.isfinite(rer))
    assert mer[-1] == rep.merit
    assert eU < TOL, eU
    assert ev < TOL, ev


def test_mixed_unique_reactions_single_layer(dev):
    """One layer of spheres on a plane, apart from each other, sliding or sticking
    with random horizontal velocities: each sphere has one contact and R is
    unique per contact."""
    s = solfec.SOLFEC("DYNAMIC", 1e-3, "/tmp/solfec_test")
    solfec.SURFACE_MATERIAL(s, friction=0.4)
    solfec.GRAVITY(s, (0.0, 0.0, -9.81))
    mat = solfec.BULK_MATERIAL(s, density=2000.0)
    solfec.BODY(s, "OBSTACLE", solfec.PLANE((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 1), mat)
    rng = np.random.default_rng(5)
    for i in range(20):
        for j in range(20):
            r = rng.uniform(0.045, 0.055)
            body = solfec.BODY(s, "RIGID", solfec.SPHERE((0.12 * i, 0.12 * j, r - 1e-4), r, 1, 1), mat)
            solfec.INITIAL_VELOCITY(body, (rng.normal(0, 0.02), rng.normal(0, 0.02), -0.01), (0.0, 0.0, 0.0))
    s.step_begin()
    assert s.ncon == 400
    rep, rer, mer, orep, Rg, Ro, Ug, Uo, vg, vo, m64 = _both(dev, s, dict(epsilon=1e-9, maxiter=200, meritval=1e-12))
    eR = _rel(Rg, Ro)
    print(f"[mixed] single layer: iters {rep.iters}, R {eR:.2e}, sliding rows "
          f"{int(np.sum(np.hypot(Ro[:, 0], Ro[:, 1]) > 0.999 * 0.4 * Ro[:, 2]))}", flush=True)
    assert eR < TOL, eR
    assert _rel(Ug, Uo) < TOL


def test_mixed_refinement_stays_within_maxiter(dev):
    """A mixed solve that reaches maxiter: the FP32 sweeps stop refine iterations
    early, the FP64 sweeps take the rest, and the histories hold exactly
    maxiter values (the FP64 merit of each iteration included)."""
    s = lattice(1000, seed=1)
    b, c = s.bodies(), s.contacts()
    dev.bodies(b["kind"], b["rot"], b["x0"], b["jinv"], b["minv"], b["velo"], b["velo0"])
    dev.contacts(c["master"], c["slave"], c["mpnt"], c["spnt"], c["base"], c["gap"], c["friction"],
                 c["restitution"], c["cohesion"], c["R"], c["U"])
    dev.assemble(1e-3)
    rep, rer, mer = dev.solve(solfec.make_param(1e-12, 40, meritval=1e-14, precision=solfec.device.MIXED, refine=2))
    assert rep.iters == 40 and rep.error == 1          # DIVERGED at maxiter, as bgs.c:1055-1058
    assert len(rer) == len(mer) == 40
    assert np.all(rer > 0) and np.all(mer > 0)
    assert mer[-1] == rep.merit and rer[-1] == rep.rel_error
    # the 38 FP32 iterations' merits decrease overall and continue into the FP64 ones
    assert mer[-1] < mer[0]


def test_mixed_from_bodies_and_body_update(dev):
    s = lattice(512, seed=7)
    b, c = s.bodies(), s.contacts()
    dev.bodies(b["kind"], b["rot"], b["x0"], b["jinv"], b["minv"], b["velo"], b["velo0"])
    dev.contacts(c["master"], c["slave"], c["mpnt"], c["spnt"], c["base"], c["gap"], c["friction"],
                 c["restitution"], c["cohesion"], c["R"], c["U"])
    dev.assemble(1e-3)
    rep, _, _ = dev.solve(solfec.make_param(1e-6, 3000, meritval=1e-8, precision=solfec.device.MIXED))
    Rg, _ = dev.get_RU()
    velo_g, _ = dev.body_update(1e-3, len(b["kind"]))
    velo_o, _ = oracle.body_update(oracle.bodies_from(b), oracle.contacts_from(c), Rg.ravel(), 1e-3)
    assert np.abs(velo_g - velo_o).max() <= 1e-12 * np.abs(velo_o).max()
