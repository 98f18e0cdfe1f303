"""Warm-start perceptron on the GPU (bf16 matrix-core GEMMs, FP32 accumulation)
against the FP64 numpy restatement in oracle/mlp.py: forward pass, loss and
every layer's gradient, Adam training, and the device warm start applied to a
contact system.  bf16 activations and weights carry 8 significant bits, so the
forward tolerance is 3 % of the output scale and gradients are compared per
layer by relative norm error (10 %) and cosine similarity (0.99)."""
import numpy as np
import pytest

from oracle import mlp as omlp
import solfec
from solfec import learned
from scenes import lattice

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = solfec.Device(0)
    yield d
    d.close()


def _data(seed, n):
    rng = np.random.default_rng(seed)
    X = (rng.normal(size=(n, 18)) * rng.uniform(0.1, 10, 18) + rng.normal(size=18)).astype(np.float32)
    Y = np.stack([np.sin(X[:, 0] / X[:, 0].std()) * 3.0, X[:, 1] * X[:, 2] / (X[:, 1].std() * X[:, 2].std()),
                  np.tanh(X[:, 3] / X[:, 3].std())], 1).astype(np.float32)
    return X, Y


def _net(seed):
    p = learned.init_params(seed)
    rng = np.random.default_rng(seed + 7)
    W = learned.unpack(p)
    for k in (1, 3, 5, 7):                      # non-zero biases so every epilogue term is exercised
        W[k][:] = rng.normal(scale=0.1, size=W[k].shape)
    W[6][:, :3] = rng.normal(scale=0.2, size=(128, 3))
    return p


@pytest.mark.parametrize("n", [1, 77, 4096])
def test_forward_matches_oracle(dev, n):
    print(f"[mlp] forward n={n}", flush=True)
    X, Y = _data(1, max(n, 64))
    X, Y = X[:n], Y[:n]
    p = _net(0)
    norm = learned.norm_from_data(*_data(1, 64))
    dev.mlp_set(p, norm)
    out = dev.mlp_forward(X)
    ref = omlp.predict(p.astype(np.float64), norm.astype(np.float64), X.astype(np.float64))
    scale = np.abs(ref).max()
    err = np.abs(out - ref).max() / scale
    print(f"[mlp] forward n={n} max rel err {err:.3e}")
    assert err < 3e-2
    # a wrong row / column placement in the matrix-core tiles gives O(1) errors; check per output too
    for c in range(3):
        assert np.corrcoef(out[:, c], ref[:, c])[0, 1] > 0.999 if n > 8 else True


def test_gradient_matches_oracle(dev):
    X, Y = _data(2, 1000)
    p = _net(3)
    norm = learned.norm_from_data(X, Y)
    dev.mlp_set(p, norm)
    loss, g = dev.mlp_grad(X, Y)
    rloss, rg = omlp.loss_and_grad(p.astype(np.float64), norm.astype(np.float64), X.astype(np.float64),
                                   Y.astype(np.float64))
    print(f"[mlp] loss gpu {loss:.6e} oracle {rloss:.6e}")
    assert abs(loss - rloss) <= 3e-2 * rloss
    offs = np.cumsum([0] + [int(np.prod(s)) for s in omlp.SHAPES])
    for k in range(8):
        a, b = g[offs[k]:offs[k + 1]].astype(np.float64), rg[offs[k]:offs[k + 1]]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
        print(f"[mlp] layer part {k}: rel {rel:.3e} cos {cos:.6f}")
        assert rel < 0.1 and cos > 0.99
    # padded inputs and outputs get zero gradient
    W1g = g[offs[0]:offs[1]].reshape(32, 128)
    assert np.all(W1g[18:] == 0)
    W4g = g[offs[6]:offs[7]].reshape(128, 32)
    assert np.all(W4g[:, 3:] == 0)


def test_training_reduces_loss(dev):
    X, Y = _data(4, 24000)            # one distribution: the last 4000 samples are held out
    X, Y, Xv, Yv = X[:20000], Y[:20000], X[20000:], Y[20000:]
    norm = learned.norm_from_data(X, Y)
    p0 = learned.init_params(11)
    dev.mlp_set(p0, norm)
    l0, _ = dev.mlp_grad(Xv, Yv)
    hist = dev.mlp_train(X, Y, epochs=6, batch=512, lr=2e-3, seed=5)
    l1, _ = dev.mlp_grad(Xv, Yv)
    print(f"[mlp] held-out loss {l0:.4e} -> {l1:.4e}; epochs {hist}")
    assert np.all(np.isfinite(hist))
    assert l1 < 0.2 * l0
    # the trained parameters reproduce on the oracle
    p1 = dev.mlp_get()
    ref = omlp.predict(p1.astype(np.float64), norm.astype(np.float64), Xv[:500].astype(np.float64))
    out = dev.mlp_forward(Xv[:500])
    assert np.abs(out - ref).max() <= 3e-2 * np.abs(ref).max()


def test_training_deterministic(dev):
    X, Y = _data(6, 3000)
    norm = learned.norm_from_data(X, Y)
    res = []
    for _ in range(2):
        dev.mlp_set(learned.init_params(2), norm)
        dev.mlp_train(X, Y, epochs=2, batch=256, lr=1e-3, seed=9)
        res.append(dev.mlp_get())
    assert np.array_equal(res[0], res[1])


def test_warm_start_on_system(dev):
    """SFGS_Warm_Start shifts each contact's R by the network's prediction for its features."""
    s = lattice(512, seed=3)
    b, c = s.bodies(), s.contacts()
    n = len(c["gap"])
    rng = np.random.default_rng(0)
    R0 = rng.normal(scale=1e-3, size=3 * n)
    dev.bodies(b["kind"], b["rot"], b["x0"], b["jinv"], b["minv"], b["velo"], b["velo0"])
    dev.contacts(c["master"], c["slave"], c["mpnt"], c["spnt"], c["base"], c["gap"], c["friction"],
                 c["restitution"], c["cohesion"], R0, c["U"])
    dev.assemble(1e-3)
    F = dev.features()
    assert F.shape == (n, 18)
    np.testing.assert_allclose(F[:, 13:16], R0.reshape(n, 3).astype(np.float32), rtol=1e-6)
    p = _net(4)
    norm = learned.norm_from_data(F, np.ones((n, 3), np.float32))
    dev.mlp_set(p, norm)
    pred = dev.mlp_forward(F)
    dev.warm_start()
    R, _ = dev.get_RU()
    np.testing.assert_allclose(R.reshape(n, 3), R0.reshape(n, 3) + pred, rtol=1e-5, atol=1e-6 * np.abs(pred).max())
