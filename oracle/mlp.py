"""ORACLE — test infrastructure only.

FP64 numpy restatement of the learned warm start of the optimal-transport
solver (no counterpart in geogram): the per-seed features, the MLP
6 -> 64 -> 64 -> 1 with leaky ReLU (slope 0.01), and the gradient of its mean
squared error by hand-written back-propagation.  Only ``tests/`` may import
this module.  Pinned by ``tests/test_warmstart.py`` against central finite
differences of ``forward`` and against torch autograd on the CPU.

Parameter vector layout (4673 values): W1[6][64], b1[64], W2[64][64] (row k =
input unit k), b2[64], W3[64], b3.
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID = 6, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_IN * N_HID
OFF_W2 = OFF_B1 + N_HID
OFF_B2 = OFF_W2 + N_HID * N_HID
OFF_W3 = OFF_B2 + N_HID
OFF_B3 = OFF_W3 + N_HID
N_PARAMS = OFF_B3 + 1
LEAK = 0.01


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    W1 = p[OFF_W1:OFF_B1].reshape(N_IN, N_HID)
    b1 = p[OFF_B1:OFF_W2]
    W2 = p[OFF_W2:OFF_B2].reshape(N_HID, N_HID)
    b2 = p[OFF_B2:OFF_W3]
    W3 = p[OFF_W3:OFF_B3]
    b3 = p[OFF_B3]
    return W1, b1, W2, b2, W3, b3


def _leaky(a):
    return np.where(a > 0, a, LEAK * a)


def _leaky_d(a):
    return np.where(a > 0, 1.0, LEAK)


def forward(p, X):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    a1 = X @ W1 + b1
    h1 = _leaky(a1)
    a2 = h1 @ W2 + b2
    h2 = _leaky(a2)
    return h2 @ W3 + b3


def loss_and_grad(p, X, T):
    """Mean over rows of (MLP(x) - t)^2 and its gradient (chain rule written
    out layer by layer)."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    n = X.shape[0]
    a1 = X @ W1 + b1
    h1 = _leaky(a1)
    a2 = h1 @ W2 + b2
    h2 = _leaky(a2)
    y = h2 @ W3 + b3
    e = y - T
    loss = float(np.mean(e * e))
    dy = 2.0 * e / n                      # d loss / d y
    g = np.zeros(N_PARAMS)
    g[OFF_W3:OFF_B3] = h2.T @ dy
    g[OFF_B3] = dy.sum()
    dh2 = np.outer(dy, W3) * _leaky_d(a2)  # d loss / d a2
    g[OFF_W2:OFF_B2] = (h1.T @ dh2).ravel()
    g[OFF_B2:OFF_W3] = dh2.sum(axis=0)
    dh1 = (dh2 @ W2.T) * _leaky_d(a1)      # d loss / d a1
    g[OFF_W1:OFF_B1] = (X.T @ dh1).ravel()
    g[OFF_B1:OFF_W2] = dh1.sum(axis=0)
    return loss, g


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, Algorithm 1); returns new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def features(xyz, nu, box, knn_d2):
    """Seed features: position scaled to the box (3), log(nu_i n / V), and the
    log of the mean and of the max distance to the 16 nearest other seeds
    over h = (V / n)^(1/3).  knn_d2: n x 16 squared distances, ascending."""
    xyz = np.asarray(xyz, dtype=np.float64)
    lo = np.asarray(box[:3], dtype=np.float64)
    hi = np.asarray(box[3:], dtype=np.float64)
    n = xyz.shape[0]
    V = float(np.prod(hi - lo))
    h = (V / n) ** (1.0 / 3.0)
    d = np.sqrt(np.asarray(knn_d2, dtype=np.float64)[:, :16])
    f = np.empty((n, 6))
    f[:, :3] = (xyz - lo) / (hi - lo)
    f[:, 3] = np.log(np.asarray(nu) * n / V)
    f[:, 4] = np.log(d.mean(axis=1) / h)
    f[:, 5] = np.log(d.max(axis=1) / h)
    return f
