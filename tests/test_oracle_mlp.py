"""Pins the FP64 MLP reference (oracle/mlp_ref.py) by central finite
differences, so it can then check the GPU gradients.  CPU only."""
import numpy as np

from oracle import mlp_ref


def fd_grad(p, x, dy, in_dim, hid, eps=1e-6):
    g = np.zeros_like(p)
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = eps
        yp, _ = mlp_ref.forward(p + e, x, in_dim, hid)
        ym, _ = mlp_ref.forward(p - e, x, in_dim, hid)
        g[i] = (dy @ yp - dy @ ym) / (2 * eps)
    return g


def make(in_dim, hid, seed):
    rng = np.random.default_rng(seed)
    n = mlp_ref.num_params(in_dim, hid)
    p = rng.normal(0, 0.5, n)
    w = mlp_ref.unpack(p, in_dim, hid)
    p[-(2 * in_dim + 2):] = np.concatenate([rng.normal(0, 1, in_dim), rng.uniform(0.5, 2, in_dim), [1.0, 0.0]])
    x = rng.normal(0, 1, (9, in_dim))
    dy = rng.normal(0, 1, 9)
    return p, x, dy


def test_gradient_matches_finite_differences():
    for seed in range(3):
        p, x, dy = make(5, 7, seed)
        g = mlp_ref.grad(p, x, dy, 5, 7)
        fd = fd_grad(p, x, dy, 5, 7)
        trained = mlp_ref.num_params(5, 7) - (2 * 5 + 2)
        assert np.allclose(g[:trained], fd[:trained], rtol=1e-5, atol=1e-7)
        assert np.all(g[trained:] == 0)


def test_forward_shapes_and_relu():
    p, x, dy = make(5, 7, 4)
    y, (xs, z1, h1, z2, h2) = mlp_ref.forward(p, x, 5, 7)
    assert y.shape == (9,) and np.all(h1 >= 0) and np.all(h2 >= 0)
    assert np.allclose(h1, np.maximum(z1, 0))
