"""ORACLE - test infrastructure only.

NumPy (float64) restatement of the per-element MLP of the Davidson preconditioner
(dmrgpp_amd/csrc/mlp_kernels.hip): 4 features -> 64 tanh -> 64 tanh -> 1, parameters packed
as W1 (64x4 row major), b1, W2 (64x64 row major), b2, W3 (64), b3. Forward pass, the mean
squared error loss and its gradient by backpropagation. There is no reference counterpart
(the reference has no Davidson solver); the gradient is pinned by central finite differences
in tests/test_davidson_gpu.py, and the GPU kernels are checked against it there.
"""
from __future__ import annotations

import numpy as np

IN, HID = 4, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + HID * IN
OFF_W2 = OFF_B1 + HID
OFF_B2 = OFF_W2 + HID * HID
OFF_W3 = OFF_B2 + HID
OFF_B3 = OFF_W3 + HID
NPARAMS = OFF_B3 + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(HID, IN), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(HID, HID),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


def forward(p, x):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    h1 = np.tanh(x @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    return h2 @ W3 + b3, h1, h2


def loss(p, x, t):
    y, _, _ = forward(p, x)
    return float(np.mean((y - np.asarray(t, dtype=np.float64)) ** 2))


def gradient(p, x, t):
    """Gradient of the mean squared error, packed like the parameters."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    y, h1, h2 = forward(p, x)
    dy = 2.0 * (y - np.asarray(t, dtype=np.float64)) / x.shape[0]
    gW3 = h2.T @ dy
    gb3 = dy.sum()
    da2 = (dy[:, None] * W3[None, :]) * (1.0 - h2 ** 2)
    gW2 = da2.T @ h1
    gb2 = da2.sum(axis=0)
    da1 = (da2 @ W2) * (1.0 - h1 ** 2)
    gW1 = da1.T @ x
    gb1 = da1.sum(axis=0)
    return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])
