"""Batched solve of independent scenes against the CPU oracle solving every
scene on its own, in the colour order the batch used for that scene and from
the same R0.  FP64: the reactions within 1e-6 of max|R| per scene, iteration
counts within one.  MIXED (FP32 sweeps, then per scene FP64 sweeps under the
reference's test, as the single-system MIXED solve): packings are
hyperstatic, so what the contact problem fixes is checked on every scene that
converges, the relative velocities U and the velocity of every body after the
step within 1e-6 of their largest magnitude, and the FP64 merit of the
returned state at or below meritval.  Also: workgroup count per scene does not
change the result, the two schedules agree (forward and alternating sweeps), a
scene without contacts, autotuning, and refusal of a split system by the
single-system entry point."""
import numpy as np
import pytest

import oracle
import solfec
from solfec import batch
from scenes import lattice

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = solfec.Device(0)
    yield d
    d.close()


def _scenes(specs):
    out = []
    for n, seed, mu in specs:
        s = lattice(n, seed=seed, friction=mu)
        out.append((s.bodies(), s.contacts()))
    return out


SPECS = [(27, 1, 0.3), (343, 2, 0.1), (64, 3, 0.8), (1000, 4, 0.5), (125, 5, 0.3), (512, 6, 0.6), (8, 7, 0.4)]


def _oracle_scene(b, c, colours, kw, R0):
    O = oracle.assemble(oracle.bodies_from(b), oracle.contacts_from(c), 1e-3)
    order, _ = oracle.colour_order(colours, int(colours.max()) + 1 if len(colours) else 1)
    n = O["n"]
    rc, rep, R, U, _, _ = oracle.gs_solve(oracle.System(O), oracle.gs_param(**kw), order, R0, np.zeros(3 * n),
                                          history=False)
    return rep, np.asarray(R).reshape(n, 3), np.asarray(U).reshape(n, 3), O


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_matches_individual_oracle(dev, precision):
    print(f"[batch] precision={precision}", flush=True)
    scenes = _scenes(SPECS)
    B, Cn, bs, cs = batch.concat(scenes)
    rng = np.random.default_rng(0)
    R0 = rng.normal(scale=1e-3, size=Cn["R"].shape)
    ns = batch.upload(dev, B, Cn, bs, 1e-3, R=R0)
    assert ns == len(SPECS)
    kw = dict(epsilon=1e-8, maxiter=4000, meritval=1e-10) if precision == 0 else \
        dict(epsilon=1e-9, maxiter=20000, meritval=1e-12)
    reps, tot = dev.batch_solve(solfec.make_param(**kw, precision=precision, refine=2), ns)
    Rg, Ug = dev.get_RU()
    vg, _ = dev.body_update(1e-3, len(B["kind"]))
    vg = np.asarray(vg).reshape(-1, 6)
    boff = np.concatenate([[0], np.cumsum([len(b["kind"]) for b, _ in scenes])])
    colours = dev.locdyn()["colour"]
    Rs, Us, Cols, R0s = (batch.split(Rg, cs, 3), batch.split(Ug, cs, 3), batch.split(colours, cs, 1),
                         batch.split(R0, cs, 3))
    checked = 0
    for k, (b, c) in enumerate(scenes):
        orep, Ro, Uo, O = _oracle_scene(b, c, Cols[k].ravel(), kw, R0s[k].ravel())
        err = _rel(Rs[k], Ro)
        print(f"[batch] scene {k}: rows {reps['rows'][k]} groups {reps['groups'][k]} iters gpu {reps['iters'][k]} "
              f"oracle {orep.iters}; merit {reps['merit'][k]:.2e}; |dR|/max|R| {err:.2e}", flush=True)
        assert reps["rows"][k] == len(c["gap"])
        if precision == 0:
            # same colour order, same R0: the batch follows the oracle iterate by iterate
            assert err < 1e-6
            assert abs(int(reps["iters"][k]) - orep.iters) <= 1
        elif orep.iters < kw["maxiter"]:
            # converged states of two precisions: what the contact problem fixes
            vo, _ = oracle.body_update(oracle.bodies_from(b), oracle.contacts_from(c), Ro.ravel(), 1e-3)
            eU = _rel(Us[k], Uo)
            ev = _rel(vg[boff[k]:boff[k + 1]].ravel(), np.asarray(vo).ravel())
            m64 = oracle.merit(oracle.System(O), Rs[k].ravel(), Us[k].ravel(), 1)[0]
            print(f"[batch]   MIXED: U {eU:.2e}, body velocities {ev:.2e}, FP64 merit of the state {m64:.2e}",
                  flush=True)
            assert reps["iters"][k] < kw["maxiter"]
            assert reps["merit"][k] <= kw["meritval"]
            assert abs(m64 - reps["merit"][k]) <= 1e-6 * m64 + 1e-18
            assert eU < 1e-6 and ev < 1e-6
            checked += 1
    if precision == 1:
        assert checked >= 4
    assert tot.iters == reps["iters"].max()
    # row updates count every sweep of both phases: an iteration is one sweep, and each
    # phase adds the sweep that closes its last iteration (two phases in MIXED)
    lo = int(np.sum(reps["iters"].astype(np.int64) * reps["rows"]))
    hi = int(np.sum((reps["iters"].astype(np.int64) + 1 + precision) * reps["rows"]))
    print(f"[batch] row updates {tot.row_updates}, iterations x rows {lo}", flush=True)
    assert lo <= tot.row_updates <= hi


def test_batch_workgroups_do_not_change_result(dev):
    scenes = _scenes(SPECS[:4])
    B, Cn, bs, cs = batch.concat(scenes)
    par = solfec.make_param(epsilon=1e-7, maxiter=2000, meritval=1e-9)
    dev.set_layout(2)
    res = []
    for tune in ((2, 4, 1 << 30), (2, 1, 32), (2, 2, 200), (2, 8, 64), (2, 4, 0)):   # lanes: the default layout
        # the solve writes R back into the uploaded contacts: upload the same R0 for every shape
        ns = batch.upload(dev, B, Cn, bs, 1e-3)
        reps, _ = dev.batch_solve(par, ns, tune=tune)
        R, _ = dev.get_RU()
        print(f"[batch] tune {tune}: groups {list(reps['groups'])} iters {list(reps['iters'])}")
        res.append((reps, R))
    for reps, R in res[1:]:
        same = reps["iters"] == res[0][0]["iters"]
        for k in range(ns):
            if same[k]:
                Ra, Rb_ = batch.split(R, cs, 3)[k], batch.split(res[0][1], cs, 3)[k]
                assert np.array_equal(Ra, Rb_)
    assert max(r["groups"].max() for r, _ in res) > 1


def test_batch_empty_scene_and_refusal(dev):
    s1 = lattice(125, seed=3)
    s2 = solfec.SOLFEC("DYNAMIC", 1e-3, "/tmp/solfec_test")
    solfec.GRAVITY(s2, (0.0, 0.0, -9.81))
    solfec.BODY(s2, "RIGID", solfec.SPHERE((0.5, 0.5, 0.5), 0.05, 1, 1), solfec.BULK_MATERIAL(s2, density=2000.0))
    s2.step_begin()
    assert s2.ncon == 0
    B, Cn, bs, cs = batch.concat([(s1.bodies(), s1.contacts()), (s2.bodies(), s2.contacts())])
    ns = batch.upload(dev, B, Cn, bs, 1e-3)
    par = solfec.make_param(epsilon=1e-6, maxiter=1000)
    reps, _ = dev.batch_solve(par, ns)
    assert reps["rows"][1] == 0 and reps["iters"][1] == 0
    assert reps["iters"][0] > 0
    with pytest.raises(solfec.SfgsError):
        dev.solve(par)
    # one scene again: the single-system entry point works
    dev.batch_scenes(np.zeros(len(B["kind"]), np.int32), nscenes=1)
    dev.assemble(1e-3)
    dev.solve(par)


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_autotune(dev, precision):
    """FP64 tries every persistent shape and the per-colour schedule for 1 and 2
    lanes per row; MIXED runs the per-colour schedule only and tunes the lanes."""
    scenes = _scenes(SPECS[:5])
    B, Cn, bs, cs = batch.concat(scenes)
    ns = batch.upload(dev, B, Cn, bs, 1e-3)
    par = solfec.make_param(epsilon=1e-6, maxiter=500, meritval=1e-6, precision=precision, refine=2)
    tune, ms = dev.batch_autotune(par, trial_iters=10)
    print(f"[batch] precision {precision}: autotuned {tune} {ms:.3f} ms")
    assert tune[0] in (1, 2) and tune[1] in (1, 2, 4, 8) and tune[2] >= 0 and tune[3] in (0, 1) and ms > 0
    if precision == 1:
        assert tune[3] == 1
    reps, _ = dev.batch_solve(par, ns)
    assert np.all(reps["merit"] <= 1e-6) or np.all(reps["iters"] > 0)
    dev.set_layout(2)


@pytest.mark.parametrize("precision,reverse", [(0, 0), (0, 1), (1, 0)])
def test_batch_schedules_agree(dev, precision, reverse):
    """The persistent launch and the per-colour schedule run the same sweeps of every
    scene, forward or alternating: bit-identical reactions whenever a scene stops
    after the same iteration (the schedules add the error and merit partials in
    different orders).  MIXED takes the per-colour schedule whatever the tune says,
    so its two runs are the same solve."""
    scenes = _scenes(SPECS)
    B, Cn, bs, cs = batch.concat(scenes)
    par = solfec.make_param(epsilon=1e-7, maxiter=2000, meritval=1e-9, precision=precision, refine=2,
                            reverse=reverse)
    dev.set_layout(2)   # the autotuner may have left another layout on the shared device
    res = []
    for tune in ((2, 4, 0, 0), (2, 4, 0, 1)):
        ns = batch.upload(dev, B, Cn, bs, 1e-3)
        reps, tot = dev.batch_solve(par, ns, tune=tune)
        R, U = dev.get_RU()
        print(f"[batch] schedule {tune[3]}: iters {list(reps['iters'])} merit {reps['merit'].max():.2e} "
              f"solve {tot.solve_ms:.2f} ms")
        res.append((reps, R, U))
    (r0, R0, U0), (r1, R1, U1) = res
    for k in range(ns):
        assert abs(int(r0["iters"][k]) - int(r1["iters"][k])) <= 1
        if r0["iters"][k] == r1["iters"][k]:
            assert np.array_equal(batch.split(R0, cs, 3)[k], batch.split(R1, cs, 3)[k])
            np.testing.assert_allclose(r1["merit"][k], r0["merit"][k], rtol=1e-9, atol=1e-300)
