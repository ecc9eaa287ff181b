"""CPU pins of the NRC oracle (oracle/nrc.py).

* backward: central finite differences of the loss (denominator held fixed,
  the stop-gradient of the relative-L2 loss) match the analytic gradient;
* relative-L2 output gradient: the reference formula's closed form for a
  one-sample case, and zero for a perfect prediction;
* bf16 rounding: exact values, ties to even, NaN;
* encoding: the nrcEncoding.ts layout, the spatial hash constants,
  trilinear corners, octahedral map and one-blob bins; the table gradient
  against finite differences;
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
    d = rng.normal(size=(n, 6))
    f[:, 3:6] = d[:, :3] / np.linalg.norm(d[:, :3], axis=1, keepdims=True)
    f[:, 6:9] = d[:, 3:] / np.linalg.norm(d[:, 3:], axis=1, keepdims=True)
    f[:, 9] = rng.random(n)
    f[:, 10:13] = rng.random((n, 3))
    return f


def test_backward_matches_finite_differences():
    w = rand_net(3)
    x = N.encode(rand_features(24), w)
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


def test_encoding_layout():
    """nrcEncoding.ts order: hash grid (16) | one-blob u (8) | one-blob v (8)
    | normal | roughness | albedo | zeros."""
    w = rand_net(2)
    f = rand_features(100)
    x = N.encode(f, w)
    assert x.shape == (100, 64)
    assert np.allclose(x[:, 16:24].sum(axis=1), 1.0) and np.allclose(x[:, 24:32].sum(axis=1), 1.0)
    assert np.all(x[:, 16:32] >= 0)
    assert np.array_equal(x[:, 32:35], f[:, 6:9].astype(np.float64))
    assert np.array_equal(x[:, 35], f[:, 9].astype(np.float64))
    assert np.array_equal(x[:, 36:39], f[:, 10:13].astype(np.float64))
    assert not x[:, 39:].any()


def test_spatial_hash_and_corners():
    """spatialHash3D constants (primes 1, 2654435761, 805459861; 32-bit wrap,
    mod 4096) and trilinear corners: weights sum to 1, a grid vertex puts all
    weight on one corner, level l has resolution 4 * 2^l."""
    assert N.grid_hash(1, 0, 0) == 1
    assert N.grid_hash(0, 1, 0) == 2654435761 % 4096
    assert N.grid_hash(0, 0, 1) == 805459861 % 4096
    assert N.grid_hash(3, 5, 7) == (3 ^ ((5 * 2654435761) & 0xFFFFFFFF) ^ ((7 * 805459861) & 0xFFFFFFFF)) % 4096
    pos = np.random.default_rng(9).random((50, 3)).astype(np.float32)
    for lv in range(N.LEVELS):
        rows, wts = N.grid_corners(lv, pos)
        assert np.allclose(wts.sum(axis=1), 1.0, atol=1e-6) and np.all(wts >= 0)
        assert rows.min() >= 0 and rows.max() < N.ROWS
    # position (0.25, 0.5, 0.75) at level 0 (resolution 4) is the vertex (1, 2, 3)
    rows, wts = N.grid_corners(0, np.array([[0.25, 0.5, 0.75]], np.float32))
    assert wts[0, 0] == 1.0 and rows[0, 0] == N.grid_hash(1, 2, 3)
    # a table holding f(row) = row interpolates to the corner's row there
    w = np.zeros(N.N_PARAMS)
    w[N.N_MLP:N.N_MLP + 2 * N.ROWS:2] = np.arange(N.ROWS)
    x = N.encode(np.array([[0.25, 0.5, 0.75, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0]], np.float32), w)
    assert x[0, 0] == N.grid_hash(1, 2, 3) and x[0, 1] == 0.0


def test_oct_encode_and_one_blob():
    u, v = N.oct_encode(np.array([[0, 0, 1], [1, 0, 0], [0, 0, -1], [0, 0, 0]], np.float64))
    assert u.tolist() == [0.5, 1.0, 1.0, 0.5] and v.tolist() == [0.5, 0.5, 1.0, 0.5]
    b = N.one_blob(np.array([0.0, 0.5, 1.0, 1.0 / 16]))
    assert np.allclose(b.sum(axis=1), 1.0)
    assert np.argmax(b[3]) == 0 and np.isclose(b[1, 3], b[1, 4])  # 0.5 sits between bins 3 and 4


def test_table_gradient_matches_finite_differences():
    """hashGridBackward: d loss / d table entry (denominator held fixed)
    against central differences of the loss through the encoding."""
    w = rand_net(13, scale=0.3)
    f = rand_features(16, seed=14)
    t = np.abs(np.random.default_rng(15).normal(0.5, 0.3, (16, 3)))
    x = N.encode(f, w)
    y, acts = N.forward_exact(w, x)
    den = N.denominators(y)
    _, d = N.relative_l2(y, t)
    W = N.split(w)
    d = (d @ W[N.HIDDEN]) * (acts[N.HIDDEN] > 0)
    for m in range(N.HIDDEN - 1, 0, -1):
        d = (d @ W[m]) * (acts[m] > 0)
    g = N.table_gradient(f, d @ W[0][:, :16])
    rows, _ = N.grid_corners(5, f[:4, 0:3])
    idx = [(5 * N.ROWS + int(r)) * 2 + c for r in rows[:, 0] for c in (0, 1)]
    h = 1e-5
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[N.N_MLP + i] += h
        wm[N.N_MLP + i] -= h
        lp = N.loss_fixed_den(wp, N.encode(f, wp), t, den)
        lm = N.loss_fixed_den(wm, N.encode(f, wm), t, den)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-9, (i, fd, g[i])
    assert np.count_nonzero(g) <= 16 * N.LEVELS * 8 * 2


def test_adam_first_step_is_signed_lr():
    w = rand_net(6)
    g = np.random.default_rng(7).normal(0, 1, w.size)
    w1, m, v = N.adam(w, g, np.zeros_like(w), np.zeros_like(w), 1)
    # m_hat = g, v_hat = g^2: the step is lr * g / (|g| + eps), lr 1e-2 for
    # the MLP and 1e-1 for the tables (nrcSubsystem.ts defaults)
    lr = np.where(np.arange(w.size) < N.N_MLP, 1e-2, 1e-1)
    assert np.allclose(w - w1, lr * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=0)
    big = np.abs(g) > 1e-3
    assert np.allclose(np.abs(w - w1)[big], lr[big], rtol=1e-5)
