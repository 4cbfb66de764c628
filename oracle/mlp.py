"""CPU restatement of the planner's cost-model network, in float64.

Test infrastructure only: the GPU network (qevedo-simulator_amd/csrc/mlp.hip)
is checked against these functions. There is no reference counterpart (the
reference plans by a fixed rule, src/blocking.rs:150-302); the gradient here
is pinned by central finite differences in tests/test_mlp_oracle.py.

Parameter layout (one flat array): W1[h, i], b1[h], W2[h, h], b2[h], W3[h],
b3 — inputs -> hidden -> hidden -> 1 with ReLU after the two hidden layers.
"""
from __future__ import annotations

import numpy as np


def num_params(inputs: int, hidden: int) -> int:
    return hidden * inputs + hidden + hidden * hidden + hidden + hidden + 1


def unpack(p: np.ndarray, inputs: int, hidden: int):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + hidden * inputs].reshape(hidden, inputs); o += hidden * inputs
    b1 = p[o:o + hidden]; o += hidden
    W2 = p[o:o + hidden * hidden].reshape(hidden, hidden); o += hidden * hidden
    b2 = p[o:o + hidden]; o += hidden
    W3 = p[o:o + hidden]; o += hidden
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


def forward(p, x, inputs: int, hidden: int):
    """y for rows of x, plus the hidden activations."""
    W1, b1, W2, b2, W3, b3 = unpack(p, inputs, hidden)
    x = np.asarray(x, dtype=np.float64).reshape(-1, inputs)
    h1 = np.maximum(x @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    y = h2 @ W3 + b3
    return y, h1, h2


def loss(p, x, t, inputs: int, hidden: int) -> float:
    y, _, _ = forward(p, x, inputs, hidden)
    return float(np.mean((y - np.asarray(t, dtype=np.float64)) ** 2))


def gradient(p, x, t, inputs: int, hidden: int):
    """d loss / d p for loss = mean over rows of (y - t)^2, by the chain rule."""
    W1, b1, W2, b2, W3, b3 = unpack(p, inputs, hidden)
    x = np.asarray(x, dtype=np.float64).reshape(-1, inputs)
    y, h1, h2 = forward(p, x, inputs, hidden)
    rows = len(x)
    dy = 2.0 * (y - np.asarray(t, dtype=np.float64)) / rows
    dW3 = h2.T @ dy
    db3 = dy.sum()
    dh2 = np.outer(dy, W3) * (h2 > 0)
    dW2 = dh2.T @ h1
    db2 = dh2.sum(axis=0)
    dh1 = (dh2 @ W2) * (h1 > 0)
    dW1 = dh1.T @ x
    db1 = dh1.sum(axis=0)
    return np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dW3, [db3]]), float(np.mean((y - t) ** 2))
