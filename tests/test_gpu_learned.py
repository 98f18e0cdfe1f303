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


@pytest.mark.parametrize("new_only,scale", [(0, 1.0), (1, 0.5)])
def test_fused_warm_start_equals_layer_by_layer_forward(dev, new_only, scale):
    """The warm start runs the whole network in one fused launch (k_mlp_infer);
    SFGS_Mlp_Forward runs it layer by layer through k_gemm.  Both form every
    product with the same bf16 MFMA operations in the same order, so the shift
    of R is the forward pass's output times the trust factor, up to the rounding
    of the last float product (the forward pass returns out times the output
    scale, the warm start multiplies scale, out and the output scale in that
    order).  With new_only, rows that carry a reaction keep it exactly."""
    from scenes import lattice
    params, norm = learned.load()
    learned.install(dev, params=params, norm=norm)
    s = lattice(1000, seed=2)
    b, c = s.bodies(), s.contacts()
    n = len(c["master"])
    R = np.random.default_rng(4).normal(scale=0.1, size=(n, 3))
    R[::3] = 0.0                             # a third of the rows start without a reaction
    dev.bodies(b["kind"], b["rot"], b["x0"], b["jinv"], b["minv"], b["velo"], b["velo0"])
    dev.contacts(c["master"], c["slave"], c["mpnt"], c["spnt"], c["base"], c["gap"], c["friction"],
                 c["restitution"], c["cohesion"], R.ravel(), c["U"])
    dev.assemble(1e-3)
    F = dev.features()
    out = dev.mlp_forward(F).reshape(n, 3).astype(np.float64)
    R0, _ = dev.get_RU()
    dev.warm_start_mode(new_only, scale)
    dev.warm_start()
    R1, _ = dev.get_RU()
    shift = R1 - R0
    want = scale * out
    carried = np.any(R0 != 0.0, axis=1)
    if new_only:
        want[carried] = 0.0
        assert np.array_equal(R1[carried], R0[carried])
    err = np.abs(shift - want).max() / np.abs(want).max()
    print(f"[learned] fused warm start, new_only={new_only} scale={scale}: {n} rows, "
          f"{int((~carried).sum())} without a reaction, max |shift - scale out| / max {err:.2e}", flush=True)
    assert np.abs(want).max() > 0
    assert err < 1e-6
    dev.warm_start_mode(1, 0.5)
