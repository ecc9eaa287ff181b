"""CPU pins of the NRC oracle (oracle/nrc.py).

* backward: central finite differences of the loss (denominator held fixed,
  the stop-gradient of the relative-L2 loss) match the analytic gradient;
* relative-L2 output gradient: the reference formula's closed form for a
  one-sample case, and zero for a perfect prediction;
* bf16 rounding: exact values, ties to even, NaN;
* encoding: ranges and the constant bias input;
* Adam: the first bias-corrected step moves every weight by lr * g / (|g| + eps).
"""
import numpy as np

from oracle import nrc as N


def rand_net(seed=0, scale=0.2):
    rng = np.random.default_rng(seed)
    return rng.normal(0, scale, N.N_PARAMS)


def rand_features(n, seed=1):
    rng = np.random.default_rng(seed)
    f = np.zeros((n, 16), np.float32)
    f[:, 0:3] = rng.random((n, 3))
    f[:, 3:7] = rng.random((n, 4))
    f[:, 7] = rng.random(n)
    f[:, 8:14] = rng.random((n, 6))
    return f


def test_backward_matches_finite_differences():
    w = rand_net(3)
    x = N.encode(rand_features(24))
    t = np.abs(np.random.default_rng(4).normal(0.5, 0.3, (24, 3)))
    y, _ = N.forward_exact(w, x)
    den = N.denominators(y)
    _, g = N.gradient_exact(w, x, t)
    rng = np.random.default_rng(5)
    idx = np.concatenate([rng.choice(16384, 40, replace=False), 16384 + rng.choice(192, 10, replace=False)])
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (N.loss_fixed_den(wp, x, t, den) - N.loss_fixed_den(wm, x, t, den)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-9, (i, fd, g[i])


def test_relative_l2_closed_form():
    y = np.array([[0.5, 0.25, 0.125]])
    t = np.array([[0.25, 0.25, 0.0]])
    lum = 0.299 * 0.5 + 0.587 * 0.25 + 0.114 * 0.125
    den = lum * lum + 1e-2
    loss, g = N.relative_l2(y, t)
    e = y - t
    assert np.isclose(loss, np.sum(e * e) / den / 3)
    assert np.allclose(g, 2 * e / 3 / den)
    loss0, g0 = N.relative_l2(t + 1.0, t + 1.0)
    assert loss0 == 0 and not g0.any()


def test_bf16_rounding():
    assert N.bf16(1.0) == 1.0
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: rounds to even (1.0)
    assert N.bf16(np.float32(1.0 + 2 ** -8)) == 1.0
    assert N.bf16(np.float32(1.0 + 3 * 2 ** -8)) == 1.0 + 2 ** -6
    assert np.isnan(N.bf16(np.nan))
    assert N.bf16(-2.5) == -2.5


def test_encoding_ranges():
    x = N.encode(rand_features(100))
    assert x.shape == (100, 64)
    assert np.all(np.abs(x[:, :36]) <= 1) and np.all((x[:, 36:52] > 0) & (x[:, 36:52] <= 1))
    assert np.all(x[:, 63] == 1) and not x[:, 59:63].any()
    # sin(pi * 2^k * p) at p = 0.25: k=0 -> sin(pi/4), k=1 -> 1, k=2 -> 0
    f = np.zeros((1, 16), np.float32)
    f[0, :3] = 0.25
    e = N.encode(f)[0]
    assert np.isclose(e[0], np.sqrt(0.5)) and np.isclose(e[1], 1.0) and abs(e[2]) < 1e-12


def test_adam_first_step_is_signed_lr():
    w = rand_net(6)
    g = np.random.default_rng(7).normal(0, 1, w.size)
    w1, m, v = N.adam(w, g, np.zeros_like(w), np.zeros_like(w), 1)
    # m_hat = g, v_hat = g^2: the step is lr * g / (|g| + eps)
    assert np.allclose(w - w1, 1e-2 * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=0)
    big = np.abs(g) > 1e-3
    assert np.allclose(np.abs(w - w1)[big], 1e-2, rtol=1e-5)
