"""The surrogate's BF16 MFMA kernels against the CPU oracle, and training.

Tolerances (normwise relative, ||gpu - oracle|| / ||oracle||):
  forward vs the bf16-rounding oracle      2e-3   (float32 vs float64 sums)
  forward vs the unrounded FP64 network    3e-2   (the cost of bf16 operands)
  gradients vs the bf16-rounding oracle    1e-2 per layer block (a bf16 rounding
      of a hidden activation can land on the other side of a tie)
  Adam step: the first step moves each parameter by lr * sign(g): >= 99 % of
      the parameters with |g| > 1e-4 max|g| move exactly as the oracle's
"""
import numpy as np
import pytest

import mlp_oracle as mo
import surrogate

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300)


def _data(E, B, seed=1):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(E, B, 4)).astype(np.float32)
    t = rng.normal(size=(E, B, 3)).astype(np.float32)
    mask = (rng.uniform(size=(E, B, 3)) > 0.2).astype(np.float32)
    return x, t, mask


def test_layout_matches_oracle():
    net = surrogate.Ensemble(2, seed=3)
    off, shp = net.layout()
    w_off, b_off, per = mo.layout()
    assert list(off[:4]) == w_off and list(off[4:]) == b_off and net.P == per
    assert list(shp) == [256, 64, 256, 256, 256, 256, 64, 256]
    p = net.get_params().reshape(2, per)
    W1 = p[:, w_off[0]:w_off[0] + 256 * 64].reshape(2, 256, 64)
    W4 = p[:, w_off[3]:w_off[3] + 64 * 256].reshape(2, 64, 256)
    assert np.all(W1[:, :, 4:] == 0) and np.all(W4[:, 3:] == 0)
    # weights and biases of the real units uniform in +-1/sqrt(fan_in), padding zero
    fan_in = [4, 256, 256, 256]
    rows = [256, 256, 256, 3]
    for l in range(4):
        W = p[:, w_off[l]:w_off[l] + mo.ROWS[l] * mo.COLS[l]].reshape(2, mo.ROWS[l], mo.COLS[l])
        bias = p[:, b_off[l]:b_off[l] + mo.ROWS[l]]
        bound = 1.0 / np.sqrt(fan_in[l])
        real = W[:, :rows[l], :fan_in[l]]
        assert np.all(np.abs(real) <= bound) and np.abs(real).max() > 0.9 * bound
        assert np.all(np.abs(bias[:, :rows[l]]) <= bound)
        assert np.all(bias[:, rows[l]:] == 0)
    assert not np.array_equal(p[0], p[1])          # members differ


def test_forward_matches_oracle():
    E, B = 3, 192
    net = surrogate.Ensemble(E, seed=7)
    p = net.get_params()
    x, _, _ = _data(E, B)
    y = net.forward(x)
    want = mo.forward(p, E, x.astype(np.float64), bf=True)
    assert _rel(y, want) <= 2e-3
    exact = mo.forward(p, E, x.astype(np.float64), bf=False)
    assert _rel(y, exact) <= 3e-2
    # a batch that is not a multiple of 64 is padded and trimmed
    y2 = net.forward(x[:, :100])
    np.testing.assert_allclose(y2, y[:, :100], rtol=0, atol=0)


def test_gradients_match_oracle():
    E, B = 2, 128
    net = surrogate.Ensemble(E, seed=11)
    p = net.get_params()
    x, t, mask = _data(E, B, seed=2)
    g, loss = net.gradients(x, t, mask)
    gw, lw = mo.gradients(p, E, x.astype(np.float64), t.astype(np.float64),
                          mask.astype(np.float64), bf=True)
    np.testing.assert_allclose(loss, lw, rtol=2e-3)
    w_off, b_off, per = mo.layout()
    G, GW = g.reshape(E, per), gw.reshape(E, per)
    for l in range(4):
        for a, n in ((w_off[l], mo.ROWS[l] * mo.COLS[l]), (b_off[l], mo.ROWS[l])):
            e = _rel(G[:, a:a + n], GW[:, a:a + n])
            assert e <= 1e-2, (l, a, e)
    # padding gradients are exactly zero
    W1 = G[:, w_off[0]:w_off[0] + 256 * 64].reshape(E, 256, 64)
    assert np.all(W1[:, :, 4:] == 0)


def test_adam_step_matches_oracle():
    E, B = 2, 64
    net = surrogate.Ensemble(E, seed=5)
    p0 = net.get_params().astype(np.float64)
    x, t, mask = _data(E, B, seed=4)
    gw, _ = mo.gradients(p0, E, x.astype(np.float64), t.astype(np.float64),
                         mask.astype(np.float64), bf=True)
    net.train_step(x, t, mask, lr=1e-3)
    p1 = net.get_params().astype(np.float64)
    want, _, _ = mo.adam(p0, np.zeros_like(p0), np.zeros_like(p0), gw, 1, 1e-3)
    big = np.abs(gw) > 1e-4 * np.abs(gw).max()
    same = np.isclose(p1[big] - p0[big], want[big] - p0[big], rtol=1e-3, atol=1e-7)
    assert same.mean() >= 0.99, same.mean()
    # parameters with zero gradient (the padding) do not move
    np.testing.assert_array_equal(p1[gw == 0], p0[gw == 0])


def test_training_fits_a_smooth_map():
    """Three smooth functions of 4 inputs: the ensemble's held-out error drops
    more than tenfold, and its spread is small where data is dense."""
    rng = np.random.default_rng(0)
    X = rng.uniform(-1, 1, size=(2000, 4))
    X[:, 3] = rng.uniform(0.5, 1.0, 2000)      # the wall-speed column (first target / 2V)
    Y = np.stack([(np.sin(2 * X[:, 0]) + X[:, 1] ** 2 + 2.0) * 2 * X[:, 3],
                  X[:, 2] * X[:, 3] + 0.5 * X[:, 0], np.tanh(X[:, 0] + X[:, 3])], 1)
    model = surrogate.Surrogate(members=5, seed=1)
    Xt, Yt = X[:1800], Y[:1800]
    Xh, Yh = X[1800:], Y[1800:]
    model.xm, model.xs = Xt.mean(0), Xt.std(0)
    Yn = model._to_net(Xt, Yt)
    model.ym, model.ys = Yn.mean(0), Yn.std(0)
    before = np.abs(model.predict(Xh)[0] - Yh).mean()
    hist = model.fit(Xt, Yt, steps=600, batch=256, lr=2e-3)
    mean, std = model.predict(Xh)
    after = np.abs(mean - Yh).mean()
    assert hist[-1] < 0.1 * hist[0]
    assert after < 0.1 * before, (before, after)
    assert np.median(std) < 0.05
