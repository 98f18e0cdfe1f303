"""FP64 numpy restatement of the warm-start perceptron (forward, loss, backward).

TEST INFRASTRUCTURE ONLY: the checker of the GPU forward pass and of the
gradients produced by the device's bf16 matrix-core kernels.  Parameter layout
and normalisation follow include/sfgs.h (SFGS_Mlp_Set): per layer [in][out]
weights then biases, inputs padded 18 -> 32, outputs 3 -> 32.

Parity: pinned in tests/test_oracle_mlp.py by central finite differences of
the loss (gradient) and by GELU against scipy's error function.
"""
import numpy as np
from scipy.special import erf

NF, NI, NH, NOP, NO = 18, 32, 128, 32, 3
SHAPES = [(NI, NH), (NH,), (NH, NH), (NH,), (NH, NH), (NH,), (NH, NOP), (NOP,)]


def unpack(params):
    out, o = [], 0
    for s in SHAPES:
        size = int(np.prod(s))
        out.append(np.asarray(params[o:o + size], np.float64).reshape(s))
        o += size
    return out


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def dgelu(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def inputs(features, norm):
    X = np.zeros((len(features), NI))
    X[:, :NF] = (np.asarray(features, np.float64) - norm[:NF]) / norm[NF:2 * NF]
    return X


def forward(params, X):
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(params)
    Z1 = X @ W1 + b1
    H1 = gelu(Z1)
    Z2 = H1 @ W2 + b2
    H2 = gelu(Z2)
    Z3 = H2 @ W3 + b3
    H3 = gelu(Z3)
    O = H3 @ W4 + b4
    return O, (X, Z1, H1, Z2, H2, Z3, H3)


def predict(params, norm, features):
    O, _ = forward(params, inputs(features, norm))
    return O[:, :NO] * norm[2 * NF:2 * NF + NO]


def loss_and_grad(params, norm, features, targets):
    """Mean squared error over the 3 outputs in normalised units and its gradient."""
    X = inputs(features, norm)
    Y = np.asarray(targets, np.float64) / norm[2 * NF:2 * NF + NO]
    O, (X, Z1, H1, Z2, H2, Z3, H3) = forward(params, X)
    B = len(X)
    D = O[:, :NO] - Y
    loss = np.mean(D * D)
    dO = np.zeros_like(O)
    dO[:, :NO] = 2.0 * D / (B * NO)
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(params)
    gW4 = H3.T @ dO
    gb4 = dO.sum(0)
    dZ3 = (dO @ W4.T) * dgelu(Z3)
    gW3 = H2.T @ dZ3
    gb3 = dZ3.sum(0)
    dZ2 = (dZ3 @ W3.T) * dgelu(Z2)
    gW2 = H1.T @ dZ2
    gb2 = dZ2.sum(0)
    dZ1 = (dZ2 @ W2.T) * dgelu(Z1)
    gW1 = X.T @ dZ1
    gb1 = dZ1.sum(0)
    grad = np.concatenate([g.ravel() for g in (gW1, gb1, gW2, gb2, gW3, gb3, gW4, gb4)])
    return loss, grad
