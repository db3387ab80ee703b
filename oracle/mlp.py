"""CPU restatement of the rank predictor (MLP 32 -> 64 -> 64 -> 1, tanh, logistic output) in
float64.  TEST INFRASTRUCTURE ONLY: the checker for the GPU forward / backward / Adam kernels
(pspaqr_public_amd/csrc/learned.hip); the product never imports it.

Parameter layout (flat, as the C ABI): W1 (64 x 32, row-major), b1 (64), W2 (64 x 64),
b2 (64), w3 (64), b3 (1).  Loss: mean_i softplus(o_i) - y_i o_i, the logistic loss with soft
targets y in [0, 1] (rank fractions).  There is no reference counterpart: the reference has
no learned component; this restates the textbook backpropagation of that network.
"""
from __future__ import annotations

import numpy as np

NF, NH = 32, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + NH * NF
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NH
NPARAM = OFF_B3 + 1


def unpack(W):
    W = np.asarray(W, dtype=np.float64)
    return (W[OFF_W1:OFF_B1].reshape(NH, NF), W[OFF_B1:OFF_W2], W[OFF_W2:OFF_B2].reshape(NH, NH), W[OFF_B2:OFF_W3],
            W[OFF_W3:OFF_B3], W[OFF_B3])


def forward(X, W):
    W1, b1, W2, b2, w3, b3 = unpack(W)
    X = np.asarray(X, dtype=np.float64)
    h1 = np.tanh(X @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    o = h2 @ w3 + b3
    return o, (X, h1, h2)


def predict(X, W):
    o, _ = forward(X, W)
    return 1.0 / (1.0 + np.exp(-o))


def loss(X, y, W):
    o, _ = forward(X, W)
    y = np.asarray(y, dtype=np.float64)
    return float(np.mean(np.maximum(o, 0) - o * y + np.log1p(np.exp(-np.abs(o)))))


def loss_grad(X, y, W):
    """(loss, flat gradient) by backpropagation."""
    W1, b1, W2, b2, w3, b3 = unpack(W)
    o, (X, h1, h2) = forward(X, W)
    y = np.asarray(y, dtype=np.float64)
    n = X.shape[0]
    L = float(np.mean(np.maximum(o, 0) - o * y + np.log1p(np.exp(-np.abs(o)))))
    do = (1.0 / (1.0 + np.exp(-o)) - y) / n
    g = np.zeros(NPARAM)
    g[OFF_W3:OFF_B3] = h2.T @ do
    g[OFF_B3] = do.sum()
    da2 = np.outer(do, w3) * (1 - h2 * h2)
    g[OFF_W2:OFF_B2] = (da2.T @ h1).ravel()
    g[OFF_B2:OFF_W3] = da2.sum(0)
    da1 = (da2 @ W2) * (1 - h1 * h1)
    g[OFF_W1:OFF_B1] = (da1.T @ X).ravel()
    g[OFF_B1:OFF_W2] = da1.sum(0)
    return L, g


def adam(W, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step (returns new W, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return W - lr * mh / (np.sqrt(vh) + eps), m, v
