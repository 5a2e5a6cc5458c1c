"""The learned depth guess: the numpy reference network pinned by finite
differences and closed forms (CPU), then the library's HIP kernels against it
(GPU): MFMA and scalar inference, forward/backward, training-set generation
and training."""
import numpy as np
import pytest

import mlp_ref as R
from conftest import load_oracle, load_trmc

trmc = load_trmc()
orc = load_oracle()


def _rand_weights(seed, scale=0.5):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(R.NPARAM) * scale).astype(np.float32)


# ------------------------------------------------------------------- CPU pins
def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    W = rng.standard_normal(R.NPARAM) * 0.4
    X = rng.uniform(-2, 2, (40, 5))
    Y = rng.standard_normal(40)
    _, g = R.loss_and_grad(W, X, Y)
    idx = np.concatenate([rng.choice(R.NPARAM, 120, replace=False), [R.OFF_B3, R.OFF_W3, R.OFF_B1, R.OFF_B2]])
    for i in idx:
        e = np.zeros(R.NPARAM)
        e[i] = 1e-6
        num = (R.loss_and_grad(W + e, X, Y)[0] - R.loss_and_grad(W - e, X, Y)[0]) / 2e-6
        assert abs(num - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, num, g[i])


def test_reference_forward_closed_forms():
    # zero hidden weights: y = b3 + sum W3 tanh(b2)
    W = np.zeros(R.NPARAM)
    W[R.OFF_B2:R.OFF_W3] = np.linspace(-1, 1, 32)
    W[R.OFF_W3:R.OFF_B3] = np.linspace(0.5, -0.5, 32)
    W[R.OFF_B3] = 0.25
    y, _, _ = R.forward(W, np.ones((3, 5)))
    assert np.allclose(y, 0.25 + np.dot(np.linspace(0.5, -0.5, 32), np.tanh(np.linspace(-1, 1, 32))))
    # one path: x0 -> h1[0] -> h2[0] -> y
    W = np.zeros(R.NPARAM)
    W[R.OFF_W1 + 0] = 2.0
    W[R.OFF_W2 + 0] = 3.0
    W[R.OFF_W3 + 0] = 5.0
    x = np.array([[0.3, 9, 9, 9, 9]])
    assert np.isclose(R.forward(W, x)[0][0], 5.0 * np.tanh(3.0 * np.tanh(0.6)))


def test_features_and_wide_depth_closed_forms():
    f = R.features(10 ** 0.5, 0.055, 25.5, 2.0, 1e-3)
    assert np.allclose(f, 0.0, atol=1e-6)
    # wide channel: Manning q = (1/n) h^(5/3) sqrt(S0) per unit width
    Q, n, bw, s0 = 50.0, 0.04, 20.0, 1e-3
    h = R.wide_depth(Q, n, bw, s0)
    assert np.isclose(bw * h ** (5 / 3) * np.sqrt(s0) / n, Q, rtol=1e-12)


def test_init_weights_glorot_and_deterministic():
    W = trmc.mlp_init_weights(5)
    assert np.array_equal(W, trmc.mlp_init_weights(5))
    assert not np.array_equal(W, trmc.mlp_init_weights(6))
    W1, b1, W2, b2, W3, b3 = R.unpack(W)
    assert np.all(b1 == 0) and np.all(b2 == 0) and b3 == 0
    assert np.abs(W1).max() <= np.sqrt(6 / 37) and np.abs(W2).max() <= np.sqrt(6 / 64)
    assert np.abs(W3).max() <= np.sqrt(6 / 33)
    assert np.abs(W2).max() > 0.8 * np.sqrt(6 / 64)


# ------------------------------------------------------------------- GPU
@pytest.mark.gpu
@pytest.mark.parametrize("n", [64, 1000, 4097])
def test_predict_mfma_and_scalar_match_reference(gpu, n):
    """MFMA wave forward (the routing kernel's) and scalar forward against the
    float64 reference; tolerance 1e-4 absolute (FP32 accumulation and the
    1 - 2/(1 + e^2x) tanh)."""
    W = _rand_weights(11 + n)
    X = np.random.default_rng(n).uniform(-2.5, 2.5, (n, 5)).astype(np.float32)
    want = R.forward(W, X)[0]
    got_m = trmc.mlp_predict(W, X, use_mfma=True)
    got_s = trmc.mlp_predict(W, X, use_mfma=False)
    print(f"n={n}: mfma max err {np.abs(got_m - want).max():.2e}, scalar {np.abs(got_s - want).max():.2e}",
          flush=True)
    assert np.abs(got_s - want).max() <= 1e-4 * (1 + np.abs(want).max())
    assert np.abs(got_m - want).max() <= 1e-4 * (1 + np.abs(want).max())
    assert np.abs(got_m - got_s).max() <= 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("pairs", [1, 40, 1000])
def test_packed_pairs_match_reference(gpu, pairs):
    """The routing kernel's packing: wanting lanes' sample pairs gathered two
    per lane pair into as few forwards as possible, results shuffled back."""
    W = _rand_weights(31 + pairs)
    X = np.random.default_rng(pairs).uniform(-2.5, 2.5, (2 * pairs, 5)).astype(np.float32)
    got = trmc.mlp_predict(W, X, use_mfma=2)
    want = R.forward(W, X)[0]
    lane = np.repeat(np.arange(pairs), 2)
    on = lane % 3 != 1
    assert np.all(np.isnan(got[~on]))
    assert np.abs(got[on] - want[on]).max() <= 1e-4 * (1 + np.abs(want).max())


@pytest.mark.gpu
def test_forward_backward_matches_reference(gpu):
    W = _rand_weights(21, 0.3)
    rng = np.random.default_rng(22)
    X = rng.uniform(-2, 2, (512, 5)).astype(np.float32)
    Y = rng.standard_normal(512).astype(np.float32)
    pred, grad = trmc.mlp_forward_backward(W, X, Y)
    want_y = R.forward(W, X)[0]
    _, want_g = R.loss_and_grad(W, X, Y)
    assert np.abs(pred - want_y).max() <= 1e-4 * (1 + np.abs(want_y).max())
    err = np.abs(grad - want_g).max()
    print(f"grad max err {err:.2e} (max |g| {np.abs(want_g).max():.2e})", flush=True)
    assert err <= 1e-4 * np.abs(want_g).max() + 1e-7


@pytest.mark.gpu
def test_dataset_features_and_targets(gpu):
    """Features are the float32 normalisation of the raw draws; targets are
    log(h_normal / h_wide) with h_normal the oracle's FP64 normal depth."""
    X, Y, raw = trmc.mlp_make_dataset(4096, seed=9, raw=True)
    Q, dx, bw, tw, twcc, nch, ncc, cs, s0 = raw.T
    assert np.allclose(X, R.features(Q, nch, bw, cs, s0), atol=2e-6)
    f32 = np.float32
    assert Q.min() >= f32(1e-3) * (1 - 1e-7) and Q.max() <= f32(1e4)
    assert s0.min() >= f32(1e-4) and s0.max() <= f32(1e-2)
    rows = np.random.default_rng(1).choice(4096, 300, replace=False)
    for i in rows:
        p = [dx[i], bw[i], tw[i], twcc[i], nch[i], ncc[i], cs[i], s0[i]]
        h = orc.normal_depth(p, Q[i], variant="cr")
        hw = float(np.float32(np.float32(Q[i]) * np.float32(nch[i]) /
                              (np.float32(bw[i]) * np.sqrt(np.float32(s0[i])))) ** np.float32(0.6))
        assert abs(Y[i] - np.log(h / hw)) <= 2e-6 * max(1.0, abs(Y[i])), (i, Y[i], np.log(h / hw))


@pytest.mark.gpu
def test_training_reduces_loss_and_is_deterministic(gpu):
    trmc.mlp_make_dataset(1 << 15, seed=7)
    W0 = trmc.mlp_init_weights(3)
    W, loss = trmc.mlp_train(W0, epochs=20, batch=512, lr=3e-3)
    W2, loss2 = trmc.mlp_train(W0, epochs=20, batch=512, lr=3e-3)
    assert np.array_equal(W, W2) and np.array_equal(loss, loss2)
    print(f"loss {loss[0]:.4g} -> {loss[-1]:.4g}", flush=True)
    assert np.all(np.diff(loss[::5]) < 0)
    Xh, Yh, _ = trmc.mlp_make_dataset(8192, seed=99)
    rmse = float(np.sqrt(np.mean((trmc.mlp_predict(W, Xh) - Yh) ** 2)))
    rmse0 = float(np.sqrt(np.mean((trmc.mlp_predict(W0, Xh) - Yh) ** 2)))
    base = float(np.sqrt(np.mean(Yh ** 2)))          # the wide-channel depth alone (y = 0)
    print(f"held-out rmse of log depth ratio: trained {rmse:.4f}, initial {rmse0:.4f}, "
          f"wide-channel depth alone {base:.4f}", flush=True)
    # the channel inputs leave the overbank share (twcc, tw, ncc unseen) irreducible
    assert rmse < 0.75 * base and rmse < 0.5 * rmse0
