"""CPU oracle of the shift-proposal network (TEST INFRASTRUCTURE ONLY).

FP64 numpy restatement of the 3 -> 64 -> 64 -> 16 GELU perceptron that
tailor_amd/csrc/mlp.hip trains: forward pass, mean-squared-error loss, the
analytic gradient of every parameter and one Adam step (Kingma & Ba 2015,
bias-corrected, beta1 0.9, beta2 0.999, eps 1e-8).  The reference has no
counterpart; the oracle is pinned by its own checks in tests/test_oracle_mlp.py
(central finite differences of the loss, GELU closed forms, a hand-computed
Adam step).

Parameter layout (one flat vector of 5456 entries):
    W1[64][3] b1[64] W2[64][64] b2[64] W3[16][64] b3[16]
"""
from __future__ import annotations

import math

import numpy as np

N_IN, N_H, N_OUT = 3, 64, 16
SIZES = [("W1", (N_H, N_IN)), ("b1", (N_H,)), ("W2", (N_H, N_H)), ("b2", (N_H,)),
         ("W3", (N_OUT, N_H)), ("b3", (N_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)

_erf = np.vectorize(math.erf, otypes=[float])


def unpack(p):
    out, k = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[k:k + n], dtype=np.float64).reshape(shape)
        k += n
    return out


def pack(d):
    return np.concatenate([d[name].ravel() for name, _ in SIZES])


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    Phi = 0.5 * (1.0 + _erf(z / math.sqrt(2.0)))
    phi = np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)
    return Phi + z * phi


def forward(p, X):
    w = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    Z1 = X @ w["W1"].T + w["b1"]
    A1 = gelu(Z1)
    Z2 = A1 @ w["W2"].T + w["b2"]
    A2 = gelu(Z2)
    Y = A2 @ w["W3"].T + w["b3"]
    return Y, (X, Z1, A1, Z2, A2)


def loss(p, X, T):
    Y, _ = forward(p, X)
    return float(np.mean((Y - np.asarray(T, dtype=np.float64)) ** 2))


def gradients(p, X, T):
    """(loss, flat gradient) of the mean squared error."""
    w = unpack(p)
    Y, (X, Z1, A1, Z2, A2) = forward(p, X)
    T = np.asarray(T, dtype=np.float64)
    n = X.shape[0]
    dY = 2.0 * (Y - T) / (n * N_OUT)
    g = {"W3": dY.T @ A2, "b3": dY.sum(0)}
    dZ2 = (dY @ w["W3"]) * gelu_grad(Z2)
    g["W2"], g["b2"] = dZ2.T @ A1, dZ2.sum(0)
    dZ1 = (dZ2 @ w["W2"]) * gelu_grad(Z1)
    g["W1"], g["b1"] = dZ1.T @ X, dZ1.sum(0)
    return float(np.mean((Y - T) ** 2)), pack(g)


def adam_step(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam update; returns (p, m, v)."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** step)
    vh = v / (1.0 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
