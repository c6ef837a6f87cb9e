"""The network oracle pinned on the CPU: analytic gradient against central
finite differences on tiny networks, Adam against hand-computed steps, and
a planted mistake that the finite-difference check catches."""
import numpy as np
import pytest

from cgxx_amd import learned
from oracle import mlp_oracle as mo


def _net(dims, seed):
    p = learned.init_params(dims, seed).astype(np.float64)
    rng = np.random.default_rng(seed + 1)
    p += 0.1 * rng.standard_normal(p.size)  # nonzero biases too
    X = rng.standard_normal((7, dims[0]))
    T = rng.standard_normal((7, dims[3]))
    return p, X, T


def _fd(f, p, h=1e-6):
    g = np.zeros_like(p)
    for k in range(p.size):
        e = np.zeros_like(p)
        e[k] = h
        g[k] = (f(p + e) - f(p - e)) / (2 * h)
    return g


@pytest.mark.parametrize("dims", [(3, 5, 4, 2), (6, 8, 8, 9), (1, 1, 1, 1)])
def test_gradient_matches_finite_differences(dims):
    p, X, T = _net(dims, 3)
    g, L = mo.gradient(p, dims, X, T)
    assert g.size == mo.param_count(dims)
    assert L == pytest.approx(mo.loss(p, dims, X, T), rel=1e-14)
    fd = _fd(lambda q: mo.loss(q, dims, X, T), p)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-8)


def test_planted_mistake_is_caught():
    """Dropping the ReLU derivative in the first layer must fail the check."""
    dims = (3, 5, 4, 2)
    p, X, T = _net(dims, 5)
    W1, b1, W2, b2, W3, b3 = mo.unpack(p, dims)
    a1, a2, Y = mo.forward(p, dims, X)
    dy = 2.0 * (Y - T) / Y.size
    dz2 = (dy @ W3.T) * (a2 > 0)
    dz1_wrong = dz2 @ W2.T  # missing * (a1 > 0)
    g_wrong = np.concatenate([(X.T @ dz1_wrong).ravel(), dz1_wrong.sum(0), (a1.T @ dz2).ravel(),
                              dz2.sum(0), (a2.T @ dy).ravel(), dy.sum(0)])
    fd = _fd(lambda q: mo.loss(q, dims, X, T), p)
    assert (a1 == 0).any()  # some unit is off, so the mistake matters
    assert not np.allclose(g_wrong, fd, rtol=1e-5, atol=1e-8)


def test_adam_first_steps_by_hand():
    """Kingma & Ba Alg. 1: at t = 1, m_hat = g and v_hat = g^2, so the step is
    lr * g / (|g| + eps); at t = 2 with the same g it is again
    lr * g / (|g| + eps) (bias correction makes m_hat = g, v_hat = g^2)."""
    g = np.array([0.5, -2.0, 1e-3, 0.0], np.float32)
    p0 = np.array([1.0, 1.0, 1.0, 1.0], np.float32)
    lr, eps = 1e-2, 1e-8
    p1, m1, v1 = mo.adam_step(p0, np.zeros(4), np.zeros(4), g, 1, lr=lr, eps=eps)
    np.testing.assert_allclose(m1, 0.1 * g, rtol=1e-6)
    np.testing.assert_allclose(v1, 0.001 * g * g, rtol=2e-5)  # 1 - float32(0.999)
    expect1 = p0 - lr * g / (np.abs(g) + eps)
    np.testing.assert_allclose(p1, expect1, rtol=1e-6, atol=1e-9)
    p2, m2, v2 = mo.adam_step(p1, m1, v1, g, 2, lr=lr, eps=eps)
    np.testing.assert_allclose(m2, 0.19 * g, rtol=1e-5)
    np.testing.assert_allclose(p2, expect1 - lr * g / (np.abs(g) + eps), rtol=1e-5, atol=1e-8)
    assert p2[3] == p0[3]  # zero gradient: no movement


def test_adam_descends_quadratic():
    """Adam on f(p) = |p - c|^2 converges to c."""
    c = np.array([0.3, -1.2, 2.0], np.float32)
    p = np.zeros(3, np.float32)
    m = np.zeros(3, np.float32)
    v = np.zeros(3, np.float32)
    for t in range(1, 3001):
        p, m, v = mo.adam_step(p, m, v, 2 * (p - c), t, lr=1e-2)
    np.testing.assert_allclose(p, c, atol=1e-3)
