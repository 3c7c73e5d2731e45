"""Rank MLP in numpy (oracle; test infrastructure only).

The learned rank predictor has no counterpart in the reference (the reference
always runs the full adaptive ID, ``skelfmm/util.py:13-51``). This module is the
CPU statement of the network the GPU kernels implement, for their checks:

* parameters, flat FP32: ``W1[64][8], b1[64], W2[64][64], b2[64], w3[64], b3``;
* ``y = w3 . g(W2 g(W1 x + b1) + b2) + b3`` with the exact GELU
  ``g(z) = z Phi(z) = z (1 + erf(z / sqrt 2)) / 2``;
* loss = mean over the batch of ``(y - t)^2``.

``forward``/``loss_and_grad`` compute in FP64 by backpropagation written out by
hand; the tests pin ``loss_and_grad`` against central finite differences of
``forward`` (so a wrong derivative or a transposed weight fails), and the GPU
forward/backward against these.
"""

from __future__ import annotations

import numpy as np
from scipy.special import erf

NIN, NH = 8, 64
SIZES = [("W1", (NH, NIN)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("w3", (NH,)), ("b3", (1,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out, o = {}, 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        out[name] = p[o:o + k].reshape(shape)
        o += k
    return out


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(p, X):
    P = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    z1 = X @ P["W1"].T + P["b1"]
    z2 = gelu(z1) @ P["W2"].T + P["b2"]
    return gelu(z2) @ P["w3"] + P["b3"][0]


def loss_and_grad(p, X, t):
    """Mean squared error and its gradient (flat, same order as the parameters)."""
    P = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    n = X.shape[0]
    z1 = X @ P["W1"].T + P["b1"]
    h1 = gelu(z1)
    z2 = h1 @ P["W2"].T + P["b2"]
    h2 = gelu(z2)
    y = h2 @ P["w3"] + P["b3"][0]
    e = y - t
    loss = float(np.mean(e * e))
    dy = 2.0 * e / n
    g_w3 = h2.T @ dy
    g_b3 = np.array([dy.sum()])
    d2 = np.outer(dy, P["w3"]) * gelu_d(z2)
    g_W2 = d2.T @ h1
    g_b2 = d2.sum(axis=0)
    d1 = (d2 @ P["W2"]) * gelu_d(z1)
    g_W1 = d1.T @ X
    g_b1 = d1.sum(axis=0)
    grad = np.concatenate([g_W1.ravel(), g_b1, g_W2.ravel(), g_b2, g_w3, g_b3])
    return loss, grad


def adam(p, g, m1, m2, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, algorithm 1); returns new (p, m1, m2)."""
    m1 = b1 * m1 + (1 - b1) * g
    m2 = b2 * m2 + (1 - b2) * g * g
    mh = m1 / (1 - b1 ** step)
    vh = m2 / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m1, m2
