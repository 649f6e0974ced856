"""Pins of the network oracle (oracle/mlp.py): the analytic gradient against
central finite differences of the loss, GELU closed forms, and one Adam step
computed by hand."""
import math

import numpy as np

from oracle import mlp


def _setup(seed=0, n=5):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(mlp.N_PARAMS) * 0.3
    X = rng.uniform(-1, 1, (n, 3))
    T = rng.standard_normal((n, 16))
    return p, X, T


def test_gelu_closed_forms():
    z = np.array([0.0, 1.0, -1.0, 8.0, -8.0])
    g = mlp.gelu(z)
    assert g[0] == 0.0
    assert abs(g[1] - 0.5 * (1 + math.erf(1 / math.sqrt(2)))) < 1e-15
    assert abs(g[1] - g[2] - 1.0) < 1e-15          # gelu(z) - gelu(-z) = z
    assert abs(g[3] - 8.0) < 1e-12 and abs(g[4]) < 1e-12
    d = mlp.gelu_grad(z)
    assert d[0] == 0.5
    # derivative by central differences
    h = 1e-6
    zz = np.linspace(-3, 3, 13)
    fd = (mlp.gelu(zz + h) - mlp.gelu(zz - h)) / (2 * h)
    assert np.max(np.abs(fd - mlp.gelu_grad(zz))) < 1e-8


def test_gradient_matches_finite_differences():
    p, X, T = _setup()
    _, g = mlp.gradients(p, X, T)
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.choice(mlp.N_PARAMS, 40, replace=False),
                          [0, 191, 192, 255, 256, 4351, 4352, 4415, 4416, 5439, 5440, 5455]])
    h = 1e-6
    for k in idx:
        e = np.zeros_like(p)
        e[k] = h
        fd = (mlp.loss(p + e, X, T) - mlp.loss(p - e, X, T)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 + 1e-6 * abs(g[k]), (k, fd, g[k])


def test_parameter_count_and_layout():
    assert mlp.N_PARAMS == 5456
    p = np.arange(mlp.N_PARAMS, dtype=float)
    w = mlp.unpack(p)
    assert w["W1"].shape == (64, 3) and w["b3"].shape == (16,)
    assert w["W2"][0, 0] == 64 * 3 + 64
    np.testing.assert_array_equal(mlp.pack(w), p)


def test_adam_first_step_is_lr_times_sign():
    # with zero moments, step 1 moves every parameter by lr * g / (|g| + eps)
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -4.0, 1e-3])
    p1, m, v = mlp.adam_step(p, g, np.zeros(3), np.zeros(3), 1, lr=0.01)
    np.testing.assert_allclose(p1, p - 0.01 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-15)
    np.testing.assert_allclose(m, 0.1 * g)
    np.testing.assert_allclose(v, 0.001 * g * g)


def test_training_target_roundtrip():
    from tailor_amd import shift_model as sm
    w = np.array([[0.1 - 0.2j, 0.3 - 0.1j, 0.2 - 0.3j, 0.4 - 0.05j,
                   0.5 - 0.4j, 0.6 - 0.2j, 0.15 - 0.25j, 0.35 - 0.35j]])
    np.testing.assert_array_equal(sm.targets_to_omegas(sm.omegas_to_targets(w)), w)
    lead = sm.leading_order(w[0])
    assert np.all(np.diff(lead.imag) <= 0)
    x = sm.normalise_inputs([0.5, 3.0], [500, 3000], [0.3, 0.9])
    np.testing.assert_allclose(x, [[-1, -1, -1], [1, 1, 1]], atol=1e-7)
