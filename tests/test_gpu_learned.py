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
    """Held-out trajectories (seeds outside the training and validation sets, mu 0.3
    and 0.6): GS iterations to epsilon 1e-6 on every fourth system of the first 48
    steps, from the previous-step R (the reference) and from the shipped network's
    R0 in the variant chosen on the validation trajectory.  The gain is printed; the
    network may not cost more than 5 % (the plain GS runs after it, so the answer
    does not depend on it)."""
    params, norm = learned.load()
    meta = learned.load_meta()
    v = (int(meta.get("new_only", 1)), float(meta.get("trust", 0.5)))
    base = lrn = 0
    for seed, mu in ((9, 0.3), (10, 0.6)):
        res = learned.held_out_iterations(dev, params, norm, seed, mu, (v,))
        print(f"[learned] held-out seed {seed} mu {mu}: previous-step R {res['previous']} iterations, "
              f"network {res[v]}", flush=True)
        base += res["previous"]
        lrn += res[v]
    print(f"[learned] held-out total: previous-step R {base}, network {lrn}, gain x{base / max(lrn, 1):.3f} "
          f"(variant new_only={v[0]} trust={v[1]}, trained on {meta.get('samples', '?')} samples)", flush=True)
    assert lrn <= 1.05 * base
