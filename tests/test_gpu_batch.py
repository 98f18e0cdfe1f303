"""Batched solve of independent scenes in one persistent launch against the
CPU oracle solving every scene on its own, in the colour order the batch used
for that scene and from the same R0: converged reactions within 1e-6 of
max|R| per scene (FP64), within 1e-4 for MIXED (FP32 sweeps + 2 FP64
refinement sweeps, as the single-system mixed path).  Also: workgroup count
per scene does not change the result, a scene without contacts, autotuning,
and refusal of a split system by the single-system entry point."""
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
    return rep, R


@pytest.mark.parametrize("precision,tol", [(0, 1e-6), (1, 1e-4)])
def test_batch_matches_individual_oracle(dev, precision, tol):
    print(f"[batch] precision={precision}", flush=True)
    scenes = _scenes(SPECS)
    B, Cn, bs, cs = batch.concat(scenes)
    rng = np.random.default_rng(0)
    R0 = rng.normal(scale=1e-3, size=Cn["R"].shape)
    ns = batch.upload(dev, B, Cn, bs, 1e-3, R=R0)
    assert ns == len(SPECS)
    # FP32 sweeps cannot resolve a relative change below ~1e-7: MIXED runs to epsilon 1e-6 and a merit bound
    kw = dict(epsilon=1e-8, maxiter=4000, meritval=1e-10) if precision == 0 else \
        dict(epsilon=1e-6, maxiter=4000, meritval=1e-8)
    reps, tot = dev.batch_solve(solfec.make_param(**kw, precision=precision, refine=2), ns)
    Rg, _ = dev.get_RU()
    colours = dev.locdyn()["colour"]
    Rs, Cols, R0s = batch.split(Rg, cs, 3), batch.split(colours, cs, 1), batch.split(R0, cs, 3)
    for k, (b, c) in enumerate(scenes):
        orep, Ro = _oracle_scene(b, c, Cols[k].ravel(), kw, R0s[k].ravel())
        scale = max(np.abs(Ro).max(), 1e-300)
        err = np.abs(Rs[k] - Ro).max() / scale
        print(f"[batch] scene {k}: rows {reps['rows'][k]} groups {reps['groups'][k]} iters gpu {reps['iters'][k]} "
              f"oracle {orep.iters}; merit {reps['merit'][k]:.2e}; |dR|/max|R| {err:.2e}")
        assert reps["rows"][k] == len(c["gap"])
        if precision == 0:
            # same colour order, same R0: the batch follows the oracle iterate by iterate
            assert err < tol
            assert abs(int(reps["iters"][k]) - orep.iters) <= 1
        elif orep.iters < kw["maxiter"]:
            # converged states of two precisions are compared, not iterates
            assert reps["merit"][k] <= kw["meritval"]
            assert err < tol
    assert tot.iters == reps["iters"].max()


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


def test_batch_autotune(dev):
    scenes = _scenes(SPECS[:5])
    B, Cn, bs, cs = batch.concat(scenes)
    ns = batch.upload(dev, B, Cn, bs, 1e-3)
    par = solfec.make_param(epsilon=1e-6, maxiter=500, meritval=1e-6, precision=1, refine=2)
    tune, ms = dev.batch_autotune(par, trial_iters=10)
    print(f"[batch] autotuned {tune} {ms:.3f} ms")
    assert tune[0] in (1, 2) and tune[1] in (1, 2, 4, 8) and tune[2] > 0 and ms > 0
    reps, _ = dev.batch_solve(par, ns)
    assert np.all(reps["merit"] <= 1e-6) or np.all(reps["iters"] > 0)


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_schedules_agree(dev, precision):
    """The persistent launch and the per-colour schedule run the same sweeps of every
    scene: bit-identical reactions whenever a scene stops after the same iteration
    (the schedules add the error and merit partials in different orders)."""
    scenes = _scenes(SPECS)
    B, Cn, bs, cs = batch.concat(scenes)
    par = solfec.make_param(epsilon=1e-7, maxiter=2000, meritval=1e-9, precision=precision, refine=2)
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
