"""FP64 reference of the learned-bound MLP -- TEST INFRASTRUCTURE ONLY.

Forward and backward of the network the library trains on the GPU
(33 -> H -> H -> 1, ReLU; flat parameter layout W1, b1, W2, b2, w3, b3, mu,
sd, scale, offset; inputs standardised with mu / sd), written in numpy float64.
Its gradients are pinned by central finite differences in
tests/test_oracle_mlp.py.
"""
import numpy as np


def unpack(p, in_dim, hid):
    p = np.asarray(p, np.float64)
    o = 0
    out = {}
    for name, size, shape in [("W1", in_dim * hid, (in_dim, hid)), ("b1", hid, (hid,)),
                              ("W2", hid * hid, (hid, hid)), ("b2", hid, (hid,)),
                              ("w3", hid, (hid,)), ("b3", 1, ()), ("mu", in_dim, (in_dim,)),
                              ("sd", in_dim, (in_dim,)), ("scale", 1, ()), ("offset", 1, ())]:
        out[name] = p[o:o + size].reshape(shape)
        o += size
    assert o == p.size
    return out


def forward(p, x, in_dim, hid):
    w = unpack(p, in_dim, hid)
    xs = (np.asarray(x, np.float64) - w["mu"]) / w["sd"]
    z1 = xs @ w["W1"] + w["b1"]
    h1 = np.maximum(z1, 0)
    z2 = h1 @ w["W2"] + w["b2"]
    h2 = np.maximum(z2, 0)
    y = h2 @ w["w3"] + w["b3"]
    return y, (xs, z1, h1, z2, h2)


def grad(p, x, dy, in_dim, hid):
    """d(sum dy * y)/d(params), flat; zero for mu, sd, scale, offset."""
    w = unpack(p, in_dim, hid)
    _, (xs, z1, h1, z2, h2) = forward(p, x, in_dim, hid)
    dy = np.asarray(dy, np.float64)
    gw3 = h2.T @ dy
    gb3 = dy.sum()
    dz2 = np.outer(dy, w["w3"]) * (z2 > 0)
    gW2 = h1.T @ dz2
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ w["W2"].T) * (z1 > 0)
    gW1 = xs.T @ dz1
    gb1 = dz1.sum(0)
    return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3],
                           np.zeros(2 * in_dim + 2)])


def num_params(in_dim, hid):
    return in_dim * hid + hid + hid * hid + hid + hid + 1 + 2 * in_dim + 2
