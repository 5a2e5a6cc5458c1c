"""Float64 numpy restatement of the 5 -> 32 -> 32 -> 1 tanh depth-guess
network (weight layout, features, forward, loss gradient).  Test
infrastructure: the GPU kernels are checked against it, and it is itself
pinned by finite differences (tests/test_mlp.py)."""
import numpy as np

NPARAM = 1281
OFF_W1, OFF_B1, OFF_W2, OFF_B2, OFF_W3, OFF_B3 = 0, 160, 192, 1216, 1248, 1280


def unpack(W):
    W = np.asarray(W, np.float64)
    return (W[OFF_W1:OFF_B1].reshape(32, 5), W[OFF_B1:OFF_W2], W[OFF_W2:OFF_B2].reshape(32, 32),
            W[OFF_B2:OFF_W3], W[OFF_W3:OFF_B3], W[OFF_B3])


def forward(W, X):
    W1, b1, W2, b2, W3, b3 = unpack(W)
    X = np.asarray(X, np.float64)
    a1 = np.tanh(X @ W1.T + b1)
    a2 = np.tanh(a1 @ W2.T + b2)
    return a2 @ W3 + b3, a1, a2


def loss_and_grad(W, X, Y):
    """mean 0.5 (y - Y)^2 and its gradient in the flat layout."""
    W1, b1, W2, b2, W3, b3 = unpack(W)
    X = np.asarray(X, np.float64)
    y, a1, a2 = forward(W, X)
    n = len(X)
    e = y - np.asarray(Y, np.float64)
    dy = e / n
    d2 = np.outer(dy, W3) * (1 - a2 ** 2)
    d1 = (d2 @ W2) * (1 - a1 ** 2)
    g = np.zeros(NPARAM)
    g[OFF_W1:OFF_B1] = (d1.T @ X).ravel()
    g[OFF_B1:OFF_W2] = d1.sum(0)
    g[OFF_W2:OFF_B2] = (d2.T @ a1).ravel()
    g[OFF_B2:OFF_W3] = d2.sum(0)
    g[OFF_W3:OFF_B3] = a2.T @ dy
    g[OFF_B3] = dy.sum()
    return 0.5 * float(np.mean(e ** 2)), g


def features(Q, n, bw, cs, s0):
    """Normalised inputs, in float32 as the kernels compute them."""
    f = np.float32
    Q, n, bw, cs, s0 = (np.asarray(v, np.float32) for v in (Q, n, bw, cs, s0))
    return np.stack([(np.log10(Q) - f(0.5)) * f(0.5), (n - f(0.055)) * f(40.0),
                     (bw - f(25.5)) * (f(1.0) / f(24.5)), cs - f(2.0), np.log10(s0) + f(3.0)], -1)


def wide_depth(Q, n, bw, s0):
    """(Q n / (bw sqrt S0))^(3/5) in float64."""
    return (np.asarray(Q, np.float64) * n / (bw * np.sqrt(s0))) ** 0.6
