"""RUN's control paths with the GPU solve (lng.c:8973-9025, sol.c:745-806,
bgs.c:1052-1066): a GAUSS_SEIDEL_SOLVER that ends at maxiter under
failure='EXIT' makes RUN raise the reference's DIVERGED error; under
'CALLBACK' the callable receives (solver, *data) and a zero return ends the
run with an error, a non-zero one lets it continue; a file named STOP in the
output directory ends RUN after the current step."""
import os

import pytest

import solfec

pytestmark = pytest.mark.gpu


def scene(outdir):
    s = solfec.SOLFEC("DYNAMIC", 1e-3, str(outdir))
    solfec.SURFACE_MATERIAL(s, friction=0.3)
    solfec.GRAVITY(s, (0.0, 0.0, -9.81))
    solfec.LATTICE(s, 125, 0.45, 0.045, 0.055, 0.005, 2000.0, 3)   # overlapping neighbours: contacts at once
    return s


def test_failure_exit_raises_diverged(tmp_path):
    s = scene(tmp_path)
    gs = solfec.GAUSS_SEIDEL_SOLVER(1e-12, 2, failure="EXIT")
    with pytest.raises(RuntimeError, match="DIVERGED"):
        solfec.RUN(s, gs, 5e-3)
    assert gs.error == "DIVERGED" and gs.iters == 2
    assert s.time == 0.0            # the failing step did not end


def test_failure_callback_continue_and_stop(tmp_path):
    calls = []

    def keep_going(solver, tag):
        calls.append((solver.iters, tag))
        return 1

    s = scene(tmp_path)
    gs = solfec.GAUSS_SEIDEL_SOLVER(1e-12, 3, failure="CALLBACK", callback=keep_going, data=("run",))
    solfec.RUN(s, gs, 3e-3)
    assert abs(s.time - 3e-3) < 1e-12
    assert calls == [(3, "run")] * 3           # once per step: every step ends at maxiter

    def stop(solver):
        calls.append("stop")
        return 0

    t = scene(tmp_path)
    gs2 = solfec.GAUSS_SEIDEL_SOLVER(1e-12, 3, failure="CALLBACK", callback=stop)
    with pytest.raises(RuntimeError, match="callback requested exit"):
        solfec.RUN(t, gs2, 3e-3)
    assert calls[-1] == "stop" and t.time == 0.0


def test_stop_file_ends_run(tmp_path):
    s = scene(tmp_path)
    gs = solfec.GAUSS_SEIDEL_SOLVER(1e-4, 100)
    solfec.RUN(s, gs, 2e-3)
    assert abs(s.time - 2e-3) < 1e-12
    open(os.path.join(tmp_path, "STOP"), "w").close()
    solfec.RUN(s, gs, 10e-3)                   # one step, then the STOP file ends the run
    assert abs(s.time - 3e-3) < 1e-12
    os.remove(os.path.join(tmp_path, "STOP"))
    solfec.RUN(s, gs, 2e-3)
    assert abs(s.time - 5e-3) < 1e-12
