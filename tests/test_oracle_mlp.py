"""The MLP oracle's gradients against central finite differences (CPU)."""

import numpy as np

from oracle import mlp


def test_gradient_matches_finite_differences(rng):
    F, H, B = 6, 5, 9
    p = mlp.init_params(F, H, rng).astype(np.float64)
    p += 0.1 * rng.normal(size=p.size)  # non-zero biases, mixed ReLU states
    x = rng.normal(size=(B, F))
    t = rng.normal(size=B)
    _, g = mlp.loss_and_grad(p, x, t, F, H)
    h = 1e-6
    fd = np.empty_like(p)
    for k in range(p.size):
        e = np.zeros_like(p)
        e[k] = h
        fd[k] = (mlp.loss_and_grad(p + e, x, t, F, H)[0] - mlp.loss_and_grad(p - e, x, t, F, H)[0]) / (2 * h)
    assert np.abs(g - fd).max() < 1e-6 * max(1.0, np.abs(fd).max())


def test_layout_and_forward_closed_form(rng):
    F, H = 3, 4
    assert mlp.param_count(F, H) == H * F + H + H * H + H + H + 1
    p = np.zeros(mlp.param_count(F, H))
    p[-1] = 0.75  # b3 only: constant output
    _, _, y = mlp.forward(p, rng.normal(size=(5, F)), F, H)
    assert np.allclose(y, 0.75)


def test_adam_first_step_is_sign_step():
    p = np.array([1.0, -2.0, 3.0])
    g = np.array([0.5, -4.0, 1e-3])
    p1, m, v = mlp.adam(p, g, np.zeros(3), np.zeros(3), 0.1, 0.9, 0.999, 1e-12, 1)
    # with bias correction the first step is lr * sign(g)
    assert np.allclose(p1, p - 0.1 * np.sign(g), atol=1e-9)
