"""Pins the numpy checker of the learned root proposal (oracle/mlp.py) against central
finite differences of its own loss.  CPU only."""
import numpy as np

from oracle import mlp


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.standard_normal(mlp.NP) * 0.3
    x = rng.standard_normal((37, 6))
    y = rng.standard_normal(37)
    loss, g = mlp.loss_grad(p, x, y)
    for k in rng.choice(mlp.NP, 60, replace=False).tolist() + [mlp.B3, mlp.W1, mlp.B1, mlp.W3]:
        e = np.zeros(mlp.NP)
        h = 1e-6
        e[k] = h
        fd = (mlp.loss_grad(p + e, x, y)[0] - mlp.loss_grad(p - e, x, y)[0]) / (2 * h)
        assert abs(fd - g[k]) < 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_gelu_closed_form():
    x = np.array([-3.0, -0.5, 0.0, 0.7, 2.0])
    np.testing.assert_allclose(mlp.gelu(0.0), 0.0)
    h = 1e-6
    np.testing.assert_allclose(mlp.gelu_grad(x), (mlp.gelu(x + h) - mlp.gelu(x - h)) / (2 * h), rtol=1e-7)


def test_param_count():
    assert mlp.NP == 4673  # 6*64 + 64 + 64*64 + 64 + 64 + 1
