"""Pins the CPU oracle of the learned warm start (oracle/mlp.py) against finite differences and
closed forms, so that the GPU kernels are checked against something independently fixed."""
import numpy as np

from oracle import mlp as OM


def _case(seed=0, n=12):
    rng = np.random.default_rng(seed)
    p = OM.init(seed).astype(np.float64)
    p[OM.OFF_B1:OM.OFF_W2] = rng.normal(0, 0.3, OM.NH)   # non-zero biases too
    p[OM.OFF_B2:OM.OFF_W3] = rng.normal(0, 0.3, OM.NH)
    p[OM.OFF_B3] = 0.2
    X = rng.normal(0, 1, (n, OM.NIN))
    y = rng.normal(0, 1, n)
    return p, X, y


def test_gradient_matches_central_differences():
    p, X, y = _case()
    g = OM.grad(p, X, y)
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.choice(OM.NPARAM, 300, replace=False),
                          np.arange(OM.OFF_W3, OM.NPARAM)])      # plus the whole output layer
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (OM.loss(p + e, X, y) - OM.loss(p - e, X, y)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_gradient_special_cases():
    p, X, y = _case(2)
    # a perfect fit has zero gradient
    assert np.abs(OM.grad(p, X, OM.forward(p, X))).max() == 0.0
    # d/db3 of the mean squared error is 2 mean(residual)
    g = OM.grad(p, X, y)
    assert np.isclose(g[OM.OFF_B3], 2 * np.mean(OM.forward(p, X) - y))


def test_forward_closed_forms():
    X = np.random.default_rng(3).normal(0, 1, (5, OM.NIN))
    p = np.zeros(OM.NPARAM)
    p[OM.OFF_B3] = 0.7
    assert np.allclose(OM.forward(p, X), 0.7)                  # zero weights -> b3
    # one hidden path: out = w3_0 tanh(W2_00 tanh(W1_0 . x + b1_0) + b2_0) + b3
    p[OM.OFF_W1:OM.OFF_W1 + OM.NIN] = np.arange(1, OM.NIN + 1) * 0.1
    p[OM.OFF_B1] = -0.2
    p[OM.OFF_W2] = 1.5
    p[OM.OFF_B2] = 0.05
    p[OM.OFF_W3] = -2.0
    a1 = X @ (np.arange(1, OM.NIN + 1) * 0.1) - 0.2
    ref = -2.0 * np.tanh(1.5 * np.tanh(a1) + 0.05) + 0.7
    # the other 63 hidden units of layer 2 see tanh(b2 = 0) = 0 with w3 = 0
    assert np.allclose(OM.forward(p, X), ref)


def test_adam_first_step_and_pack_roundtrip():
    p, X, y = _case(4)
    g = OM.grad(p, X, y)
    p1, m, v = OM.adam_step(p, g, np.zeros_like(p), np.zeros_like(p), 1, lr=1e-3, eps=0.0)
    nz = g != 0
    # bias-corrected first step moves every parameter by lr in the direction of -sign(g)
    assert np.allclose((p - p1)[nz], 1e-3 * np.sign(g[nz]))
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)
    assert np.array_equal(OM.pack(*OM.unpack(p)), p)
    assert OM.NPARAM == 4737


def test_features_definition():
    rng = np.random.default_rng(5)
    n = 50
    u, up = rng.uniform(0, 1, n), rng.uniform(0, 1, n)
    phi, gp = rng.normal(0, 0.1, n), rng.uniform(0, 6, n)
    bc = (rng.uniform(0, 1, n) < 0.2).astype(np.uint8)
    X, sigma = OM.features(u, up, phi, gp, rng.uniform(0, 3, n), 5.3, 0.005, 0.01, bc)
    assert sigma == np.abs(u - up).max()
    assert np.abs(X[:, 1]).max() == 1.0
    assert X[:, 2].min() >= -1 and X[:, 2].max() <= 1 and X[:, 3].max() <= 1
    assert np.allclose(X[:, 4], 0.5 * np.log(5.3))
    assert np.array_equal(X[:, 6], bc)
