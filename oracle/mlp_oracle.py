"""CPU oracle of the FP32 network kernels (TEST INFRASTRUCTURE ONLY).

Restates, in numpy, what include/cghip_mlp.h computes: a network
in -> h1 -> h2 -> out with ReLU after the hidden layers, the mean squared
error L = sum (Y - T)^2 / (rows * out), its analytic gradient by
back-propagation, and one Adam step (Kingma & Ba, "Adam: A Method for
Stochastic Optimization", ICLR 2015, Algorithm 1) in the form the kernel
evaluates it: m_hat = m * corr1, v_hat = v * corr2 with corr = 1 / (1 - beta^t)
formed in FP64 from the float beta and rounded to float32.  No reference counterpart: the reference has no learned
component.  The analytic gradient is pinned by central finite differences
in tests, the Adam step by hand-computed steps.

Parameter layout: W1[in][h1] b1[h1] W2[h1][h2] b2[h2] W3[h2][out] b3[out].
"""
from __future__ import annotations

import numpy as np


def param_count(dims) -> int:
    i, h1, h2, o = dims
    return i * h1 + h1 + h1 * h2 + h2 + h2 * o + o


def unpack(params, dims):
    i, h1, h2, o = dims
    p = np.asarray(params)
    k = 0
    out = []
    for shape in ((i, h1), (h1,), (h1, h2), (h2,), (h2, o), (o,)):
        size = int(np.prod(shape))
        out.append(p[k:k + size].reshape(shape))
        k += size
    return out


def forward(params, dims, X, dtype=np.float64):
    W1, b1, W2, b2, W3, b3 = (a.astype(dtype) for a in unpack(params, dims))
    X = np.asarray(X, dtype=dtype)
    a1 = np.maximum(X @ W1 + b1, 0)
    a2 = np.maximum(a1 @ W2 + b2, 0)
    return a1, a2, a2 @ W3 + b3


def loss(params, dims, X, T) -> float:
    _, _, Y = forward(params, dims, X)
    return float(np.mean((Y - np.asarray(T, np.float64)) ** 2))


def gradient(params, dims, X, T):
    """(flat gradient, loss) in float64."""
    W1, b1, W2, b2, W3, b3 = (a.astype(np.float64) for a in unpack(params, dims))
    X = np.asarray(X, np.float64)
    T = np.asarray(T, np.float64)
    a1, a2, Y = forward(params, dims, X)
    rows, out = Y.shape
    dy = 2.0 * (Y - T) / (rows * out)
    dz2 = (dy @ W3.T) * (a2 > 0)
    dz1 = (dz2 @ W2.T) * (a1 > 0)
    g = [X.T @ dz1, dz1.sum(0), a1.T @ dz2, dz2.sum(0), a2.T @ dy, dy.sum(0)]
    return np.concatenate([a.ravel() for a in g]), float(np.mean((Y - T) ** 2))


def adam_step(params, m, v, grad, t, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step in float32 with the kernel's expression order.

    Returns (params, m, v) as new float32 arrays."""
    f = np.float32
    p = np.asarray(params, f).copy()
    g = np.asarray(grad, f)
    b1, b2, lr_, eps_ = f(beta1), f(beta2), f(lr), f(eps)
    # the library receives beta as float and forms the corrections in FP64
    corr1 = f(1.0 / (1.0 - float(np.float32(beta1)) ** t))
    corr2 = f(1.0 / (1.0 - float(np.float32(beta2)) ** t))
    m = b1 * np.asarray(m, f) + (f(1) - b1) * g
    v = b2 * np.asarray(v, f) + (f(1) - b2) * (g * g)
    mhat = m * corr1
    vhat = v * corr2
    p = p - lr_ * mhat / (np.sqrt(vhat) + eps_)
    return p.astype(f), m.astype(f), v.astype(f)

