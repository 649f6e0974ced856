"""The shift-proposal network's kernels against the FP64 oracle
(oracle/mlp.py): forward pass, loss, every parameter's gradient, Adam steps,
and that training fits a smooth target."""
import numpy as np
import pytest

from oracle import mlp as om

pytestmark = pytest.mark.gpu


def _net(seed=2):
    from tailor_amd.shift_model import Network
    return Network(0, seed)


def test_forward_and_gradients_match_oracle():
    net = _net()
    p = net.params.astype(np.float64)
    rng = np.random.default_rng(3)
    X = rng.uniform(-1, 1, (300, 3)).astype(np.float32)
    T = rng.standard_normal((300, 16)).astype(np.float32)
    Y = net.predict(X)
    Y_ref, _ = om.forward(p, X.astype(np.float64))
    assert np.max(np.abs(Y - Y_ref)) <= 1e-5 * max(1.0, np.max(np.abs(Y_ref)))
    loss, g = net.gradients(X, T)
    loss_ref, g_ref = om.gradients(p, X.astype(np.float64), T.astype(np.float64))
    print(f"loss {loss:.6e} vs {loss_ref:.6e}; max |g - g_ref| {np.max(np.abs(g - g_ref)):.2e}, "
          f"max |g_ref| {np.max(np.abs(g_ref)):.2e}", flush=True)
    assert abs(loss - loss_ref) <= 1e-5 * loss_ref
    assert np.max(np.abs(g - g_ref)) <= 2e-5 * np.max(np.abs(g_ref))
    net.close()


def test_adam_steps_match_oracle():
    net = _net(5)
    p = net.params.astype(np.float64)
    rng = np.random.default_rng(4)
    X = rng.uniform(-1, 1, (64, 3)).astype(np.float32)
    T = rng.standard_normal((64, 16)).astype(np.float32)
    hist = net.train(X, T, steps=5, lr=1e-2)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    losses = []
    for step in range(1, 6):
        l, g = om.gradients(p, X.astype(np.float64), T.astype(np.float64))
        losses.append(l)
        p, m, v = om.adam_step(p, g, m, v, step, 1e-2)
    got = net.params.astype(np.float64)
    print(f"after 5 Adam steps: max |p - p_ref| {np.max(np.abs(got - p)):.2e}", flush=True)
    np.testing.assert_allclose(hist, losses, rtol=1e-4)
    assert np.max(np.abs(got - p)) <= 1e-4
    net.close()


def test_training_fits_a_smooth_map():
    net = _net(7)
    rng = np.random.default_rng(8)
    X = rng.uniform(-1, 1, (512, 3)).astype(np.float32)
    k = np.arange(16)
    T = (np.sin(X[:, :1] * (1 + k / 8)) + X[:, 1:2] * X[:, 2:3] * np.cos(k)).astype(np.float32)
    hist = net.train(X, T, steps=1500, lr=3e-3)
    print(f"loss {hist[0]:.3e} -> {hist[-1]:.3e}", flush=True)
    assert hist[-1] < 0.02 * hist[0]
    Xt = rng.uniform(-1, 1, (128, 3)).astype(np.float32)
    Tt = np.sin(Xt[:, :1] * (1 + k / 8)) + Xt[:, 1:2] * Xt[:, 2:3] * np.cos(k)
    assert np.mean((net.predict(Xt) - Tt) ** 2) < 0.05 * np.mean(Tt ** 2)
    net.close()
