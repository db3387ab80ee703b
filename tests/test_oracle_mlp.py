"""The rank predictor's CPU oracle (oracle/mlp.py) pinned by finite differences, a closed
form for the zero network and the closed form of Adam's first step."""
import numpy as np

from oracle import mlp


def _data(n=16, seed=0):
    rng = np.random.RandomState(seed)
    X = rng.uniform(0, 1, size=(n, mlp.NF))
    y = rng.uniform(0, 1, size=n)
    W = rng.normal(0, 0.3, size=mlp.NPARAM)
    return X, y, W


def test_gradient_matches_central_differences():
    X, y, W = _data()
    _, g = mlp.loss_grad(X, y, W)
    h = 1e-6
    idx = np.r_[0:40, mlp.OFF_B1:mlp.OFF_B1 + 8, mlp.OFF_W2:mlp.OFF_W2 + 40, mlp.OFF_B2:mlp.OFF_B2 + 8,
                mlp.OFF_W3:mlp.NPARAM]
    fd = np.empty(len(idx))
    for t, p in enumerate(idx):
        Wp, Wm = W.copy(), W.copy()
        Wp[p] += h
        Wm[p] -= h
        fd[t] = (mlp.loss(X, y, Wp) - mlp.loss(X, y, Wm)) / (2 * h)
    assert np.max(np.abs(fd - g[idx])) <= 1e-7 * max(1.0, np.max(np.abs(fd)))


def test_zero_network_closed_form():
    X, y, _ = _data(n=10, seed=1)
    W = np.zeros(mlp.NPARAM)
    W[mlp.OFF_B3] = 0.7
    L, g = mlp.loss_grad(X, y, W)
    s = 1 / (1 + np.exp(-0.7))
    assert abs(L - (np.log1p(np.exp(0.7)) - 0.7 * y.mean())) < 1e-12
    assert abs(g[mlp.OFF_B3] - (s - y.mean())) < 1e-12
    assert np.all(g[:mlp.OFF_B3] == 0)  # tanh(0) = 0 kills every other path
    assert np.allclose(mlp.predict(X, W), s)


def test_adam_first_step():
    rng = np.random.RandomState(2)
    W = rng.normal(size=50)
    g = rng.normal(size=50)
    W1, m, v = mlp.adam(W, g, np.zeros(50), np.zeros(50), 1, 1e-2)
    assert np.allclose(W1, W - 1e-2 * g / (np.abs(g) + 1e-8))
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)
