"""The CPU oracle of the learned-bracket network: analytic gradients against
central finite differences, SiLU against its closed form, Adam against a hand
iteration, and a fit that lowers the loss."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as M  # noqa: E402


def test_layout():
    assert M.NPARAM == 64 * 6 + 64 + 64 * 64 + 64 + 2 * 64 + 2 == 4738


def test_silu_closed_form():
    z = np.linspace(-20, 20, 2001)
    assert np.allclose(M.silu(z), z / (1 + np.exp(-z)), rtol=1e-14, atol=1e-300)
    h = 1e-6
    fd = (M.silu(z + h) - M.silu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - M.dsilu(z))) < 1e-8
    assert M.dsilu(np.array([0.0]))[0] == 0.5


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    theta = M.init_params(3).astype(np.float64)
    x = rng.normal(size=(37, 6))
    t = rng.normal(size=(37, 2))
    loss, g = M.loss_and_grad(theta, x, t)
    idx = np.concatenate([rng.choice(M.NPARAM, 200, replace=False),
                          [0, M.OFF_B1, M.OFF_W2, M.OFF_B2, M.OFF_W3, M.OFF_B3, M.NPARAM - 1]])
    h = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (M.loss_and_grad(tp, x, t)[0] - M.loss_and_grad(tm, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_adam_matches_hand_iteration():
    theta = np.array([1.0, -2.0])
    g = np.array([0.5, -0.25])
    m = np.zeros(2)
    v = np.zeros(2)
    th, m1, v1 = M.adam_step(theta, g, m, v, 1, 0.1)
    # first step: m_hat = g, v_hat = g^2 -> update = lr * sign(g) (up to eps)
    assert np.allclose(th, theta - 0.1 * np.sign(g), atol=1e-6)
    assert np.allclose(m1, 0.1 * g) and np.allclose(v1, 0.001 * g * g)


def test_fit_lowers_loss():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(512, 6))
    t = np.stack([np.tanh(x[:, 0] + 0.5 * x[:, 1]), x[:, 2] * x[:, 3] * 0.3], 1)
    theta = M.init_params(5).astype(np.float64)
    m = np.zeros_like(theta)
    v = np.zeros_like(theta)
    l0 = M.loss_and_grad(theta, x, t)[0]
    for it in range(1, 301):
        loss, g = M.loss_and_grad(theta, x, t)
        theta, m, v = M.adam_step(theta, g, m, v, it, 3e-3)
    assert loss < 0.1 * l0
