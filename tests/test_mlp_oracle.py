"""The CPU oracle of the planner's cost-model network: its gradient against
central finite differences of its own loss, and its forward pass on a case
small enough to work out by hand."""
import numpy as np

from oracle import mlp


def test_forward_by_hand():
    # 2 inputs, 2 hidden: W1 = I, b1 = (0, -1); W2 = [[1, 1], [0, 2]], b2 = 0;
    # W3 = (1, -1), b3 = 0.5
    p = np.array([1, 0, 0, 1, 0, -1, 1, 1, 0, 2, 0, 0, 1, -1, 0.5], dtype=np.float64)
    assert p.size == mlp.num_params(2, 2)
    y, h1, h2 = mlp.forward(p, [[3.0, 0.5], [-1.0, 4.0]], 2, 2)
    # row 0: h1 = relu(3, -0.5) = (3, 0); h2 = (3, 0); y = 3 + 0.5
    # row 1: h1 = relu(-1, 3) = (0, 3); h2 = (3, 6); y = 3 - 6 + 0.5
    assert np.allclose(h1, [[3, 0], [0, 3]])
    assert np.allclose(h2, [[3, 0], [3, 6]])
    assert np.allclose(y, [3.5, -2.5])


def test_gradient_matches_finite_differences():
    rng = np.random.Generator(np.random.PCG64(3))
    inputs, hidden, rows = 5, 7, 11
    p = rng.normal(0, 0.5, mlp.num_params(inputs, hidden))
    x = rng.normal(0, 1, (rows, inputs))
    t = rng.normal(0, 1, rows)
    g, loss = mlp.gradient(p, x, t, inputs, hidden)
    assert abs(loss - mlp.loss(p, x, t, inputs, hidden)) < 1e-12
    h = 1e-6
    fd = np.empty_like(p)
    for k in range(p.size):
        e = np.zeros_like(p)
        e[k] = h
        fd[k] = (mlp.loss(p + e, x, t, inputs, hidden) - mlp.loss(p - e, x, t, inputs, hidden)) / (2 * h)
    assert np.max(np.abs(g - fd)) < 1e-6 * max(1.0, np.max(np.abs(fd)))


def test_planted_mistake_is_caught():
    # dropping the ReLU mask from the second layer's back-propagation must
    # change the gradient where a unit is inactive
    rng = np.random.Generator(np.random.PCG64(4))
    inputs, hidden = 3, 4
    p = rng.normal(0, 1, mlp.num_params(inputs, hidden))
    x = rng.normal(0, 1, (6, inputs))
    t = rng.normal(0, 1, 6)
    W1, b1, W2, b2, W3, b3 = mlp.unpack(p, inputs, hidden)
    y, h1, h2 = mlp.forward(p, x, inputs, hidden)
    dy = 2 * (y - t) / len(x)
    wrong_dh2 = np.outer(dy, W3)  # mask dropped
    right_dh2 = np.outer(dy, W3) * (h2 > 0)
    assert (h2 == 0).any() and not np.allclose(wrong_dh2.T @ h1, right_dh2.T @ h1)
