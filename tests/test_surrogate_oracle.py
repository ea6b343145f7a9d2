"""Pins the surrogate's CPU oracle against mathematics that does not depend on it.

  back-propagation   central finite differences of the FP64 loss, for parameters
                     of every layer (weights and biases) and every output
  padding            gradients of the zero padding are exactly zero
  bf16 rounding      round-to-nearest-even against exact known values
  Adam               first step moves every parameter by lr * sign(g)
"""
import numpy as np
import pytest

import mlp_oracle as mo


def _params(members, seed=0, scale=0.3):
    w_off, b_off, per = mo.layout()
    rng = np.random.default_rng(seed)
    p = np.zeros((members, per))
    real_rows = [256, 256, 256, 3]
    real_cols = [4, 256, 256, 256]
    for e in range(members):
        for l in range(4):
            W = np.zeros((mo.ROWS[l], mo.COLS[l]))
            W[:real_rows[l], :real_cols[l]] = rng.normal(0, scale / np.sqrt(real_cols[l]),
                                                         (real_rows[l], real_cols[l])) * 2
            p[e, w_off[l]:w_off[l] + W.size] = W.ravel()
            p[e, b_off[l]:b_off[l] + real_rows[l]] = rng.normal(0, 0.1, real_rows[l])
    return p.ravel()


def _batch(members, B, seed=1):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(members, B, 4))
    t = rng.normal(size=(members, B, 3))
    mask = (rng.uniform(size=(members, B, 3)) > 0.2).astype(np.float64)
    return x, t, mask


def test_backprop_matches_finite_differences():
    E, B = 2, 8
    p = _params(E)
    x, t, mask = _batch(E, B)
    g, _ = mo.gradients(p, E, x, t, mask, bf=False)
    w_off, b_off, per = mo.layout()
    rng = np.random.default_rng(5)
    picks = []
    for e in range(E):
        for l in range(4):
            rows = 3 if l == 3 else 256
            cols = 4 if l == 0 else 256
            for _ in range(3):
                r, c = rng.integers(rows), rng.integers(cols)
                picks.append(e * per + w_off[l] + r * mo.COLS[l] + c)
            picks.append(e * per + b_off[l] + rng.integers(rows))
    for k in picks:
        h = 1e-5
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        e = k // per
        lp = mo.gradients(pp, E, x, t, mask, bf=False)[1][e]
        lm = mo.gradients(pm, E, x, t, mask, bf=False)[1][e]
        fd = (lp - lm) / (2 * h)
        assert g[k] == pytest.approx(fd, rel=1e-5, abs=1e-9), k


def test_padding_gradients_are_zero():
    E, B = 1, 16
    p = _params(E)
    x, t, mask = _batch(E, B)
    for bf in (False, True):
        g, _ = mo.gradients(p, E, x, t, mask, bf=bf)
        w_off, b_off, per = mo.layout()
        W1 = g[w_off[0]:w_off[0] + 256 * 64].reshape(256, 64)
        W4 = g[w_off[3]:w_off[3] + 64 * 256].reshape(64, 256)
        assert np.all(W1[:, 4:] == 0) and np.all(W4[3:] == 0)
        assert np.all(g[b_off[3] + 3:b_off[3] + 64] == 0)


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-3],
                    np.float32)
    got = mo.bf16(vals)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even -> 1; 1 + 3*2^-8 -> 1 + 2^-6
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0 + 2 ** -7
    assert got[3] == 1.0 + 2 ** -6 and got[4] == -2.5
    assert abs(got[5] - 3e-3) <= 3e-3 * 2 ** -8


def test_adam_first_step_is_sign_step():
    rng = np.random.default_rng(3)
    p = rng.normal(size=100)
    g = rng.normal(size=100)
    p1, m, v = mo.adam(p, np.zeros(100), np.zeros(100), g, step=1, lr=1e-3)
    np.testing.assert_allclose(p1 - p, -1e-3 * np.sign(g), rtol=1e-4)   # eps = 1e-8
    np.testing.assert_allclose(m, 0.1 * g)
    np.testing.assert_allclose(v, 0.001 * g * g)
