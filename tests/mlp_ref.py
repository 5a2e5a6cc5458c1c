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


def draw_training_set(n, seed):
    """Training pairs drawn as the library's dataset kernel draws them (Q
    log-uniform [1e-3, 1e4], channel parameters from the workload
    distributions, float-representable); returns (Q, params [n, 8])."""
    rng = np.random.default_rng(seed)
    f32 = lambda x: np.asarray(x, np.float32).astype(np.float64)
    Q = f32(10 ** (-3 + 7 * rng.random(n)))
    dx, nch, bw = f32(500 + 4500 * rng.random(n)), f32(0.03 + 0.05 * rng.random(n)), f32(1 + 49 * rng.random(n))
    cs = f32(1 + 2 * rng.random(n))
    tw = f32(bw + 2 + 18 * rng.random(n))
    twcc = f32(tw + 5 + 95 * rng.random(n))
    ncc = f32(nch * (1.2 + 0.8 * rng.random(n)))
    s0 = f32(np.clip(np.exp(np.log(1e-3) + rng.standard_normal(n)), 1e-4, 1e-2))
    return Q, np.stack([dx, bw, tw, twcc, nch, ncc, cs, s0], 1)


def train_adam(X, Y, epochs=20, batch=1024, lr=3e-3, seed=3):
    """Float64 Adam with the library's schedule (x0.3 after 60 % and 85 %)."""
    rng = np.random.default_rng(seed)
    W = np.zeros(NPARAM)
    a1, a2, a3 = np.sqrt(6 / 37), np.sqrt(6 / 64), np.sqrt(6 / 33)
    W[OFF_W1:OFF_B1] = rng.uniform(-a1, a1, 160)
    W[OFF_W2:OFF_B2] = rng.uniform(-a2, a2, 1024)
    W[OFF_W3:OFF_B3] = rng.uniform(-a3, a3, 32)
    m, v, step = np.zeros(NPARAM), np.zeros(NPARAM), 0
    for ep in range(epochs):
        f = ep / epochs
        lre = lr * (1.0 if f < 0.6 else 0.3 if f < 0.85 else 0.09)
        for b in range(len(X) // batch):
            sl = slice(b * batch, (b + 1) * batch)
            _, g = loss_and_grad(W, X[sl], Y[sl])
            step += 1
            m = 0.9 * m + 0.1 * g
            v = 0.999 * v + 0.001 * g * g
            W -= lre * (m / (1 - 0.9 ** step)) / (np.sqrt(v / (1 - 0.999 ** step)) + 1e-8)
    return W


def rating_log_depth(W, Q, params):
    """log of the network's rating curve h(Q) = h_wide(Q) e^{y(Q)}."""
    Q = np.maximum(np.atleast_1d(np.asarray(Q, np.float64)), 1e-6)
    p = np.atleast_2d(params)
    x = features(Q, p[:, 4], p[:, 1], p[:, 6], p[:, 7])
    return np.log(wide_depth(Q.astype(np.float32).astype(np.float64), p[:, 4], p[:, 1], p[:, 7])) + forward(W, x)[0]
