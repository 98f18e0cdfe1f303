"""The learned warm start end to end on the GPU: snapshots written by the solver
itself (SOLFEC_Capture on config-1 trajectories), training by the library's
own bf16 matrix-core kernels, and what the shipped network changes in the
Gauss-Seidel iteration count against the previous-step reaction, on a
held-out trajectory.  The network is applied to new contacts only (R_prev = 0)
with trust factor 0.5, the variant that measured best; the gain is printed,
and the test bounds the loss it may cause (the plain GS runs after it in every
case, so the answer does not depend on the network)."""
import numpy as np
import pytest

import solfec
from solfec import learned

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = solfec.Device(0)
    yield d
    d.close()


def test_snapshots_and_training(dev):
    X, Y = learned.snapshots(11, 0.4, steps=40, epsilon=1e-6, maxiter=2000)
    assert X.shape[1] == 18 and Y.shape == (len(X), 3) and len(X) > 1000
    assert np.all(np.isfinite(X)) and np.all(np.isfinite(Y))
    new = np.all(X[:, 13:16] == 0, axis=1)
    assert new.any() and not new.all()
    params, norm, hist = learned.train(dev, X[new], Y[new], epochs=20, batch=512, lr=2e-3, seed=3)
    print(f"[learned] {new.sum()} new-contact samples, loss {hist[0]:.4f} -> {hist[-1]:.4f}")
    assert hist[-1] < hist[0]


def test_shipped_model_against_previous_reaction(dev):
    params, norm = learned.load()
    s = solfec.SOLFEC("DYNAMIC", 1e-3, "/tmp/solfec_test")
    solfec.SURFACE_MATERIAL(s, friction=0.3, restitution=0.0)
    solfec.GRAVITY(s, (0.0, 0.0, -9.81))
    solfec.LATTICE(s, 1000, 1.0, 0.045, 0.055, 0.005, 2000.0, 9)   # a seed outside the training set
    gs = solfec.GAUSS_SEIDEL_SOLVER(1e-6, 3000)
    par = solfec.make_param(1e-6, 3000, meritval=1.0)
    base = lrn = 0
    for k in range(48):
        s.step_begin()
        if k % 4 == 3:
            b, c = s.bodies(), s.contacts()
            its = []
            for use in (False, True):
                dev.bodies(b["kind"], b["rot"], b["x0"], b["jinv"], b["minv"], b["velo"], b["velo0"])
                dev.contacts(c["master"], c["slave"], c["mpnt"], c["spnt"], c["base"], c["gap"], c["friction"],
                             c["restitution"], c["cohesion"], c["R"], c["U"])
                dev.assemble(1e-3)
                if use:
                    dev.mlp_set(params, norm)
                    dev.warm_start_mode(1, 0.5)
                    dev.warm_start()
                rep, _, _ = dev.solve(par, history=False)
                its.append(rep.iters)
            base += its[0]
            lrn += its[1]
        s.step_solve(gs)
        s.step_end()
    gain = base / max(lrn, 1)
    print(f"[learned] held-out: previous-step R {base} iterations, learned {lrn}; gain x{gain:.3f}")
    assert lrn <= 1.10 * base
