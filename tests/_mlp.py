"""The precision selector's MLP restated in FP64 numpy: the checker of the HIP kernels (test infrastructure).
Layout of the flat parameters: W1 (32 × 8 row-major), b1 (32), W2 (32 × 32), b2 (32), W3 (32), b3 (1)."""

import numpy as np

F, H = 8, 32
P = H * F + H + H * H + H + H + 1


def unpack(theta):
    theta = np.asarray(theta, dtype=np.float64)
    o = 0
    W1 = theta[o:o + H * F].reshape(H, F); o += H * F
    b1 = theta[o:o + H]; o += H
    W2 = theta[o:o + H * H].reshape(H, H); o += H * H
    b2 = theta[o:o + H]; o += H
    W3 = theta[o:o + H]; o += H
    b3 = theta[o]
    return W1, b1, W2, b2, W3, b3


def forward(theta, x):
    W1, b1, W2, b2, W3, b3 = unpack(theta)
    a1 = np.tanh(x @ W1.T + b1)
    a2 = np.tanh(a1 @ W2.T + b2)
    z = a2 @ W3 + b3
    return 1.0 / (1.0 + np.exp(-z)), z, a1, a2


def loss(theta, x, y):
    _, z, _, _ = forward(theta, x)
    return float(np.mean(y * np.logaddexp(0.0, -z) + (1.0 - y) * np.logaddexp(0.0, z)))


def gradient(theta, x, y):
    W1, b1, W2, b2, W3, b3 = unpack(theta)
    p, z, a1, a2 = forward(theta, x)
    d3 = (p - y) / x.shape[0]
    d2 = np.outer(d3, W3) * (1.0 - a2 ** 2)
    d1 = (d2 @ W2) * (1.0 - a1 ** 2)
    return np.concatenate([(d1.T @ x).reshape(-1), d1.sum(0), (d2.T @ a1).reshape(-1), d2.sum(0), d3 @ a2,
                           [d3.sum()]])
