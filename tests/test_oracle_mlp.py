"""Pins the FP64 selector-network oracle (oracle/mlp.py) by central finite
differences before it checks the GPU gradients."""
import numpy as np

from oracle import mlp


def test_mlp_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = mlp.random_params(rng)
    X = rng.normal(0, 2, (7, mlp.N_IN))
    T = rng.normal(0, 1, (7, mlp.N_OUT))
    V = rng.random((7, mlp.N_OUT)) < 0.7
    _, g = mlp.loss_grad(p, X, T, V)
    idx = rng.choice(mlp.n_trainable(), 60, replace=False)
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mlp.loss_grad(pp, X, T, V)[0] - mlp.loss_grad(pm, X, T, V)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_mlp_sizes():
    assert mlp.n_params() == 5540 and mlp.n_trainable() == 5508
