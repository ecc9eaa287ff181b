"""Binary64 oracle of the neural radiance cache (vitrum_amd/csrc/nrc.hip).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

Mirrors the device cache: the input encoding of
walkaround-hybrid/src/neural/nrc/nrcEncoding.ts (multiresolution hash grid
with trainable tables, one-blob octahedral direction, raw surface features),
four 64x64 ReLU layers and a 3x64 linear output (no biases), with bf16
rounding of weights, inputs, activations and deltas at exactly the points
where the device rounds (MFMA operands) and binary64 elsewhere (the grid
corners and trilinear weights in binary32, as on the device).
The loss and its output gradient follow
walkaround-hybrid/src/neural/nrc/wgsl/fusedMlp.wgsl.ts:65-120
(nrcRelativeL2OutputGradient); Adam follows
walkaround-hybrid/src/neural/nrc/fusedMlpTrainer.ts:30-84 (the tables with
their own learning rate, nrcSubsystem.ts tableLearningRate); the table
gradient is hashGridBackward (nrcEncoding.ts): d loss / d input scattered
with the trilinear weights, colliding rows accumulating.
`forward_exact`/`backward_exact` are the same network without rounding; a
finite-difference test pins them.
"""
from __future__ import annotations

import numpy as np

N_FEAT, DIMS, WIDTH, OUT, HIDDEN = 16, 64, 64, 3, 4
N_MLP = HIDDEN * WIDTH * WIDTH + OUT * WIDTH
LEVELS, ROWS, LFEAT, MIN_RES, BLOB = 8, 4096, 2, 4, 8
N_TABLE = LEVELS * ROWS * LFEAT
N_PARAMS = N_MLP + N_TABLE
LUMA = (0.299, 0.587, 0.114)
EPS = 1e-2


def bf16(x):
    """Round to bf16 (nearest even), returned as float64 values."""
    f = np.ascontiguousarray(np.asarray(x, np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, 0x7FC0, r).astype(np.uint32) << 16
    return r.view(np.float32).astype(np.float64)


def split(w):
    """Master vector -> [W0..W3 (64x64), W4 (3x64)] (row-major [out][in])."""
    w = np.asarray(w, np.float64)
    mats = [w[l * 4096:(l + 1) * 4096].reshape(64, 64) for l in range(HIDDEN)]
    mats.append(w[16384:16384 + OUT * 64].reshape(OUT, 64))
    return mats


def tables(w):
    """The hash-grid feature tables of a master vector, [level, row, feature]."""
    return np.asarray(w, np.float64)[N_MLP:N_PARAMS].reshape(LEVELS, ROWS, LFEAT)


def grid_hash(ix, iy, iz):
    """spatialHash3D (nrcEncoding.ts): XOR of coordinate x prime products
    (1, 2654435761, 805459861) with 32-bit wraparound, mod ROWS."""
    ix, iy, iz = (np.asarray(a, np.uint64) for a in (ix, iy, iz))
    h = (ix ^ ((iy * np.uint64(2654435761)) & np.uint64(0xFFFFFFFF)) ^
         ((iz * np.uint64(805459861)) & np.uint64(0xFFFFFFFF))) & np.uint64(0xFFFFFFFF)
    return (h % np.uint64(ROWS)).astype(np.int64)


def grid_corners(level, pos):
    """trilinearCorners of one level for normalised positions [n, 3]:
    (rows [n, 8], weights [n, 8]), computed in binary32."""
    p = np.asarray(pos, np.float32) * np.float32(MIN_RES << level)
    x0 = np.floor(p)
    fr = (p - x0).astype(np.float32)
    i0 = x0.astype(np.int64)
    rows = np.zeros((p.shape[0], 8), np.int64)
    wts = np.zeros((p.shape[0], 8), np.float32)
    one = np.float32(1.0)
    for k in range(8):
        c = [(k >> a) & 1 for a in range(3)]
        w = [fr[:, a] if c[a] else one - fr[:, a] for a in range(3)]
        wts[:, k] = (w[0] * w[1]).astype(np.float32) * w[2]
        rows[:, k] = grid_hash(i0[:, 0] + c[0], i0[:, 1] + c[1], i0[:, 2] + c[2])
    return rows, wts.astype(np.float64)


def oct_encode(d):
    """octEncodeDir (nrcEncoding.ts) of directions [n, 3] -> (u, v)."""
    d = np.asarray(d, np.float64)
    sc = np.max(np.abs(d), axis=1)
    ok = (sc > 0) & np.isfinite(sc)
    s = d / np.where(ok, sc, 1.0)[:, None]
    l1 = np.sum(np.abs(s), axis=1)
    inv = 1.0 / np.where(ok, l1, 1.0)
    px, py = s[:, 0] * inv, s[:, 1] * inv
    neg = s[:, 2] < 0
    ox = (1 - np.abs(py)) * np.where(px >= 0, 1.0, -1.0)
    oy = (1 - np.abs(px)) * np.where(py >= 0, 1.0, -1.0)
    px, py = np.where(neg, ox, px), np.where(neg, oy, py)
    u = np.where(ok, px * 0.5 + 0.5, 0.5)
    v = np.where(ok, py * 0.5 + 0.5, 0.5)
    return u, v


def one_blob(x):
    """oneBlobEncodeScalar: BLOB Gaussian bins at (i + 0.5)/BLOB, sigma
    1/BLOB, normalised to unit L1."""
    uc = np.clip(np.asarray(x, np.float64), 0.0, 1.0)
    c = (np.arange(BLOB) + 0.5) / BLOB
    a = np.exp(-0.5 * ((uc[:, None] - c[None, :]) * BLOB) ** 2)
    s = a.sum(axis=1, keepdims=True)
    return np.where(s > 1e-20, a / np.where(s > 1e-20, s, 1.0), a)


def encode(features, w):
    """64-dim MLP input of [n, 16] feature records (nrcEncoding.ts order:
    hash grid 16 | one-blob u 8 | one-blob v 8 | normal 3 | roughness |
    albedo 3 | zeros), binary64."""
    f = np.asarray(features, np.float32)
    n = f.shape[0]
    x = np.zeros((n, DIMS), np.float64)
    tab = tables(w)
    for lv in range(LEVELS):
        rows, wts = grid_corners(lv, f[:, 0:3])
        x[:, 2 * lv:2 * lv + 2] = np.einsum("nk,nkf->nf", wts, tab[lv][rows])
    u, v = oct_encode(f[:, 3:6])
    x[:, 16:24] = one_blob(u)
    x[:, 24:32] = one_blob(v)
    x[:, 32:35] = f[:, 6:9]
    x[:, 35] = f[:, 9]
    x[:, 36:39] = f[:, 10:13]
    return x


def table_gradient(features, dx):
    """hashGridBackward: scatter d loss / d input (dims 0..15, [n, 16]) into
    the tables with the trilinear weights; returns a flat N_TABLE vector."""
    f = np.asarray(features, np.float32)
    g = np.zeros((LEVELS, ROWS, LFEAT), np.float64)
    for lv in range(LEVELS):
        rows, wts = grid_corners(lv, f[:, 0:3])
        for k in range(8):
            for c in range(LFEAT):
                np.add.at(g[lv, :, c], rows[:, k], wts[:, k] * dx[:, 2 * lv + c])
    return g.ravel()


def forward(w, features):
    """Device-rounded forward. Returns (y [n,3], acts [H0..H4] (bf16 values))."""
    W = [bf16(m) for m in split(w)]
    h = bf16(encode(features, w))
    acts = [h]
    for l in range(HIDDEN):
        h = bf16(np.maximum(h @ W[l].T, 0.0))
        acts.append(h)
    y = h @ W[HIDDEN].T
    return y, acts


def relative_l2(y, t, n=None):
    """Loss and output gradient (nrcRelativeL2OutputGradient), binary64."""
    y = np.asarray(y, np.float64)
    t = np.asarray(t, np.float64)
    n = y.shape[0] if n is None else n
    sc = np.max(np.abs(y), axis=1)
    safe = np.where(sc > 0, sc, 1.0)
    nl = (LUMA[0] * (y[:, 0] / safe) + LUMA[1] * (y[:, 1] / safe)) + LUMA[2] * (y[:, 2] / safe)
    lum = np.where(sc > 0, np.clip(nl, -1, 1) * sc, 0.0)
    bl = np.minimum(np.abs(lum), 1e19)
    den = bl * bl + EPS
    e = y - t
    loss = float(np.sum(e * e / den[:, None]) / (3.0 * n))
    g = np.clip(e * (2.0 / n / 3.0 / den)[:, None], -65504.0, 65504.0)
    return loss, g


def gradient(w, features, targets):
    """Device-rounded backward: (loss, grad vector of N_PARAMS)."""
    W = [bf16(m) for m in split(w)]
    y, acts = forward(w, features)
    loss, dy = relative_l2(y, targets)
    d = bf16(dy)
    grads = [None] * (HIDDEN + 1)
    grads[HIDDEN] = d.T @ acts[HIDDEN]
    d = bf16((d @ W[HIDDEN]) * (acts[HIDDEN] > 0))
    for m in range(HIDDEN - 1, -1, -1):
        grads[m] = d.T @ acts[m]
        if m > 0:
            d = bf16((d @ W[m]) * (acts[m] > 0))
    dx = d @ W[0][:, :16]   # d loss / d (hash-grid inputs), from the bf16 deltas of layer 1
    return loss, np.concatenate([g.ravel() for g in grads] + [table_gradient(features, dx)])


def forward_exact(w, x):
    W = split(w)
    h = np.asarray(x, np.float64)
    acts = [h]
    for l in range(HIDDEN):
        h = np.maximum(h @ W[l].T, 0.0)
        acts.append(h)
    return h @ W[HIDDEN].T, acts


def loss_fixed_den(w, x, t, den):
    """Relative-L2 loss with the denominator held at `den` (the stop-gradient
    term), so that its derivative is the loss gradient of the model."""
    y, _ = forward_exact(w, x)
    e = y - np.asarray(t, np.float64)
    return float(np.sum(e * e / den[:, None]) / (3.0 * y.shape[0]))


def denominators(y):
    y = np.asarray(y, np.float64)
    sc = np.max(np.abs(y), axis=1)
    safe = np.where(sc > 0, sc, 1.0)
    nl = (LUMA[0] * (y[:, 0] / safe) + LUMA[1] * (y[:, 1] / safe)) + LUMA[2] * (y[:, 2] / safe)
    lum = np.where(sc > 0, np.clip(nl, -1, 1) * sc, 0.0)
    bl = np.minimum(np.abs(lum), 1e19)
    return bl * bl + EPS


def gradient_exact(w, x, t):
    """Unrounded backward with the denominator held fixed (stop gradient)."""
    W = split(w)
    y, acts = forward_exact(w, x)
    loss, d = relative_l2(y, t)
    grads = [None] * (HIDDEN + 1)
    grads[HIDDEN] = d.T @ acts[HIDDEN]
    d = (d @ W[HIDDEN]) * (acts[HIDDEN] > 0)
    for m in range(HIDDEN - 1, -1, -1):
        grads[m] = d.T @ acts[m]
        if m > 0:
            d = (d @ W[m]) * (acts[m] > 0)
    return loss, np.concatenate([g.ravel() for g in grads])


def adam(w, g, m, v, step, lr=1e-2, b1=0.9, b2=0.999, eps=1e-8, lr_table=1e-1):
    """One Adam update in binary64 (fusedMlpTrainer.ts:30-84); the tables
    (entries N_MLP.. of the vector) use lr_table."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    rate = np.full(np.shape(w), lr, np.float64)
    rate[N_MLP:] = lr_table
    w = np.clip(w - rate * mh / (np.sqrt(vh) + eps), -65504.0, 65504.0)
    return w, m, v
