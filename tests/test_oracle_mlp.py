"""Pinning of the FP64 warm-start perceptron restatement (oracle/mlp.py).

The gradient is checked against central finite differences of the loss, GELU
against its closed-form symmetry and the error function, and the layout of the
parameter vector by a network whose output is known in closed form.
"""
import numpy as np
import pytest

from oracle import mlp

NPARAM = sum(int(np.prod(s)) for s in mlp.SHAPES)


def random_net(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    p = rng.normal(scale=scale, size=NPARAM)
    norm = np.concatenate([rng.normal(size=18), rng.uniform(0.5, 2.0, 18), rng.uniform(0.5, 2.0, 3)])
    return p, norm


def test_param_count():
    assert NPARAM == 32 * 128 + 128 + 2 * (128 * 128 + 128) + 128 * 32 + 32 == 41376


def test_gelu_closed_forms():
    x = np.linspace(-6, 6, 241)
    assert mlp.gelu(0.0) == 0.0
    np.testing.assert_allclose(mlp.gelu(x) - mlp.gelu(-x), x, atol=1e-14)      # Phi(x) + Phi(-x) = 1
    np.testing.assert_allclose(mlp.gelu(np.array([1.0])), [0.8413447460685429], rtol=1e-14)  # Phi(1)
    h = 1e-6
    np.testing.assert_allclose(mlp.dgelu(x), (mlp.gelu(x + h) - mlp.gelu(x - h)) / (2 * h), atol=1e-8)


def test_layout_known_output():
    """W1 routes feature 0 to hidden unit 5, the hidden chain passes it on, W4 reads it into output 2."""
    p = np.zeros(NPARAM)
    W1, b1, W2, b2, W3, b3, W4, b4 = mlp.unpack(p)
    offs = np.cumsum([0] + [int(np.prod(s)) for s in mlp.SHAPES])
    p[offs[0] + 0 * 128 + 5] = 1.0        # W1[0][5]
    p[offs[2] + 5 * 128 + 7] = 1.0        # W2[5][7]
    p[offs[4] + 7 * 128 + 9] = 1.0        # W3[7][9]
    p[offs[6] + 9 * 32 + 2] = 1.0         # W4[9][2]
    p[offs[7] + 1] = 0.25                 # b4[1]
    norm = np.concatenate([np.zeros(18), np.ones(18), np.array([1.0, 2.0, 3.0])])
    f = np.zeros((3, 18))
    f[:, 0] = [0.5, -1.0, 2.0]
    out = mlp.predict(p, norm, f)
    g = mlp.gelu(mlp.gelu(mlp.gelu(f[:, 0])))
    np.testing.assert_allclose(out[:, 2], 3.0 * g, rtol=1e-14)
    np.testing.assert_allclose(out[:, 1], 2.0 * 0.25, rtol=1e-14)
    np.testing.assert_allclose(out[:, 0], 0.0)


@pytest.mark.parametrize("seed", [0, 1])
def test_gradient_finite_differences(seed):
    p, norm = random_net(seed)
    rng = np.random.default_rng(100 + seed)
    X = rng.normal(size=(37, 18)) * norm[18:36] + norm[:18]
    Y = rng.normal(size=(37, 3))
    loss, g = mlp.loss_and_grad(p, norm, X, Y)
    assert g.shape == (NPARAM,)
    offs = np.cumsum([0] + [int(np.prod(s)) for s in mlp.SHAPES])
    # every layer's weights and biases, including padded-input rows that must get zero gradient
    idx = []
    for k in range(8):
        idx += list(rng.integers(offs[k], offs[k + 1], 6))
    idx += [offs[0] + 20 * 128 + 3]                  # W1 row of padded input 20
    idx += [offs[6] + 4 * 32 + 10, offs[7] + 10]     # W4 / b4 columns of padded outputs
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mlp.loss_and_grad(pp, norm, X, Y)[0] - mlp.loss_and_grad(pm, norm, X, Y)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])
    assert g[offs[0] + 20 * 128 + 3] == 0.0
    assert g[offs[7] + 10] == 0.0


def test_loss_definition():
    """Mean over samples and the 3 outputs, in normalised target units."""
    p = np.zeros(NPARAM)
    norm = np.concatenate([np.zeros(18), np.ones(18), np.array([2.0, 2.0, 2.0])])
    Y = np.array([[2.0, 4.0, -2.0], [0.0, 0.0, 6.0]])
    loss, _ = mlp.loss_and_grad(p, norm, np.zeros((2, 18)), Y)
    assert loss == pytest.approx((1 + 4 + 1 + 0 + 0 + 9) / 6.0, rel=1e-15)
