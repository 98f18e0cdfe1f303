"""Mixed precision (FP32 W/B/R sweeps with FP32 accumulation; FP64 merit,
errors, termination and body update; FP64 refinement sweeps) against the FP64
oracle on the same system, same colour order, same R0.  The error of the
converged R against the FP64 result is measured and bounded."""
import numpy as np
import pytest

import oracle
import solfec
from scenes import lattice

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = solfec.Device(0)
    yield d
    d.close()


def _fp64_reference(O, kw, colours):
    order, _ = oracle.colour_order(colours, int(colours.max()) + 1)
    n = O["n"]
    rc, rep, R, U, rer, mer = oracle.gs_solve(oracle.System(O), oracle.gs_param(**kw), order,
                                              np.zeros(3 * n), np.zeros(3 * n))
    return rep, R


@pytest.mark.parametrize("n,seed", [(343, 3), (1000, 1)])
def test_mixed_reaches_merit_and_matches_fp64(dev, n, seed):
    print(f"[mixed] n={n} seed={seed}", flush=True)
    s = lattice(n, seed=seed)
    b, c = s.bodies(), s.contacts()
    O = oracle.assemble(oracle.bodies_from(b), oracle.contacts_from(c), 1e-3)
    kw = dict(epsilon=1e-6, maxiter=5000, meritval=1e-8)
    dev.system(O["rowptr"], O["col"], O["W"], O["Wd"], O["A"], O["B"], O["V"], O["rho"], O["mu"], O["rest"],
               O["coh"], O["gap"], O["free_energy"], O["step"], np.zeros(3 * O["n"]), np.zeros(3 * O["n"]))
    colours = dev.locdyn()["colour"]
    orep, Ro = _fp64_reference(O, kw, colours)
    rep, rer, mer = dev.solve(solfec.make_param(**kw, precision=solfec.device.MIXED, refine=2))
    Rg, Ug = dev.get_RU()
    scale = np.abs(Ro).max()
    err = np.abs(Rg - Ro).max() / scale
    print(f"[mixed] iters fp64={orep.iters} mixed={rep.iters} merit={rep.merit:.3e} |dR|/max|R|={err:.3e}")
    assert rep.merit <= kw["meritval"]
    assert len(rer) == rep.iters and np.all(np.isfinite(rer))
    assert err < 1e-4
    # the final merit is the FP64 merit of the returned state
    m64 = oracle.merit(oracle.System(O), Rg.ravel(), Ug.ravel(), 1)[0]
    assert abs(m64 - rep.merit) <= 1e-6 * max(m64, 1e-300) + 1e-15


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
