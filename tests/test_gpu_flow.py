"""The persistent schedules of the sweeps -- dataflow (rows wait for their own
neighbours) and barrier (a device-wide barrier per colour) -- against the
per-colour schedule on the same system:
every row reads the values the serial colour-major sweep reads, so the
reactions and velocities are bit-identical whenever both stop after the same
number of iterations (the error and merit sums are formed in a different order,
so a stop can move by one iteration at the threshold).  Also against the CPU
oracle within 1e-6 of max|R|."""
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


def _system(dev, O, R0):
    n = O["n"]
    dev.system(O["rowptr"], O["col"], O["W"], O["Wd"], O["A"], O["B"], O["V"], O["rho"], O["mu"], O["rest"],
               O["coh"], O["gap"], O["free_energy"], O["step"], R0, np.zeros(3 * n))


@pytest.mark.parametrize("n,seed,lanes,kw", [
    (343, 2, 1, dict(epsilon=1e-6, maxiter=2000)),
    (1000, 1, 1, dict(epsilon=1e-6, maxiter=3000, meritval=1e-6)),
    (1000, 3, 2, dict(epsilon=1e-6, maxiter=3000, meritval=1e-6)),
    (512, 4, 1, dict(epsilon=1e-5, maxiter=50, meritval=1e-12)),       # stops at maxiter
    (729, 5, 1, dict(epsilon=1e-6, maxiter=3000, meritval=1e-8, precision=1, refine=2)),
    (1000, 6, 2, dict(epsilon=1e-6, maxiter=3000, meritval=1e-6, precision=1, refine=2)),
])
@pytest.mark.parametrize("sched", [1, 3])
def test_flow_equals_colour_schedule(dev, n, seed, lanes, kw, sched):
    print(f"[flow] schedule {sched} n={n} seed={seed} lanes={lanes} {kw}", flush=True)
    s = lattice(n, seed=seed)
    b, c = s.bodies(), s.contacts()
    O = oracle.assemble(oracle.bodies_from(b), oracle.contacts_from(c), 1e-3)
    R0 = np.random.default_rng(seed).normal(scale=1e-3, size=3 * O["n"])
    dev.set_layout(lanes)
    out = []
    for sc in (0, sched):
        dev.set_schedule(sc)
        _system(dev, O, R0)
        rep, rer, mer = dev.solve(solfec.make_param(**kw))
        R, U = dev.get_RU()
        out.append((rep, rer, mer, R, U))
    dev.set_schedule(0)
    dev.set_layout(1)
    (r0, rer0, mer0, R0g, U0g), (r1, rer1, mer1, R1g, U1g) = out
    print(f"[flow] iters colour {r0.iters} flow {r1.iters}; merit {r0.merit:.3e} {r1.merit:.3e}; "
          f"solve ms {r0.solve_ms:.2f} {r1.solve_ms:.2f}")
    assert abs(r0.iters - r1.iters) <= 1
    print(f"[flow] max |dR| {np.abs(R0g - R1g).max():.3e}, max |dU| {np.abs(U0g - U1g).max():.3e}", flush=True)
    if r0.iters == r1.iters:
        assert np.array_equal(R0g, R1g)
        assert np.array_equal(U0g, U1g)
        np.testing.assert_allclose(rer1, rer0, rtol=1e-10, atol=0)
        np.testing.assert_allclose(mer1, mer0, rtol=1e-10, atol=0)
    scale = np.abs(R0g).max()
    assert np.abs(R1g - R0g).max() <= 1e-6 * scale


@pytest.mark.parametrize("kw", [dict(epsilon=1e-6, maxiter=3000, meritval=1e-6, reverse=1),
                                dict(epsilon=1e-6, maxiter=3000, failure=1)])
def test_barrier_schedule_reverse_and_exit(dev, kw):
    """The barrier schedule also runs alternating sweeps and FAILURE_EXIT."""
    s = lattice(1000, seed=7)
    b, c = s.bodies(), s.contacts()
    O = oracle.assemble(oracle.bodies_from(b), oracle.contacts_from(c), 1e-3)
    R0 = np.random.default_rng(7).normal(scale=1e-3, size=3 * O["n"])
    dev.set_layout(2)
    out = []
    for sc in (0, 3):
        dev.set_schedule(sc)
        _system(dev, O, R0)
        rep, rer, mer = dev.solve(solfec.make_param(**kw))
        out.append((rep, dev.get_RU()))
    dev.set_schedule(0)
    dev.set_layout(1)
    (r0, (R0g, U0g)), (r3, (R3g, U3g)) = out
    print(f"[bar] {kw}: iters {r0.iters} {r3.iters}, error {r0.error} {r3.error}", flush=True)
    assert abs(r0.iters - r3.iters) <= 1 and r0.error == r3.error
    if r0.iters == r3.iters:
        assert np.array_equal(R0g, R3g) and np.array_equal(U0g, U3g)


def test_flow_matches_oracle(dev):
    s = lattice(1000, seed=1)
    b, c = s.bodies(), s.contacts()
    O = oracle.assemble(oracle.bodies_from(b), oracle.contacts_from(c), 1e-3)
    kw = dict(epsilon=1e-6, maxiter=5000, meritval=1e-6)
    dev.set_schedule(1)
    _system(dev, O, np.zeros(3 * O["n"]))
    rep, _, _ = dev.solve(solfec.make_param(**kw))
    Rg, _ = dev.get_RU()
    colours = dev.locdyn()["colour"]
    dev.set_schedule(0)
    order, _ = oracle.colour_order(colours, int(colours.max()) + 1)
    rc, orep, Ro, Uo, _, _ = oracle.gs_solve(oracle.System(O), oracle.gs_param(**kw), order, np.zeros(3 * O["n"]),
                                             np.zeros(3 * O["n"]), history=False)
    err = np.abs(Rg.reshape(-1, 3) - Ro).max() / np.abs(Ro).max()
    print(f"[flow] oracle iters {orep.iters} gpu {rep.iters} |dR|/max|R| {err:.2e}")
    assert abs(rep.iters - orep.iters) <= 1
    assert err < 1e-6


@pytest.mark.parametrize("lanes", [1, 2])
def test_row_order_and_block_mapping_same_iterates(dev, lanes):
    """Rows of a colour are independent: ordering them by contact index, by body
    block or by the Z-order of the master body instead of by length, mapping
    workgroups to XCDs differently, or
    grouping 1, 2 or 4 slices per workgroup changes only the layout, not the
    reactions."""
    s = lattice(1000, seed=2)
    b, c = s.bodies(), s.contacts()
    kw = dict(epsilon=1e-6, maxiter=3000, meritval=1e-6)
    out = []
    dev.set_layout(lanes)
    for order, sched, wpb in ((0, 0, 4), (1, 0, 4), (2, 0, 4), (3, 0, 4), (3, 2, 4), (3, 0, 1), (3, 0, 2),
                             (3, 0, 0)):
        dev.set_row_order(order)
        dev.set_schedule(sched)
        dev.set_sweep_block(wpb)
        dev.bodies(b["kind"], b["rot"], b["x0"], b["jinv"], b["minv"], b["velo"], b["velo0"])
        dev.contacts(c["master"], c["slave"], c["mpnt"], c["spnt"], c["base"], c["gap"], c["friction"],
                     c["restitution"], c["cohesion"], c["R"], c["U"])
        dev.assemble(1e-3)
        rep, _, _ = dev.solve(solfec.make_param(**kw))
        R, U = dev.get_RU()
        out.append((rep, R, U, dev.locdyn()["colour"]))
    dev.set_row_order(0)
    dev.set_schedule(0)
    dev.set_sweep_block(0)
    dev.set_layout(1)
    r0, R0, U0, c0 = out[0]
    for r1, R1, U1, c1 in out[1:]:
        print(f"[order] lanes {lanes}: iters {r0.iters} {r1.iters}")
        assert np.array_equal(c0, c1)
        assert abs(r0.iters - r1.iters) <= 1
        if r0.iters == r1.iters:
            assert np.array_equal(R0, R1) and np.array_equal(U0, U1)
