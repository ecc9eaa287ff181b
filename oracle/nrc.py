"""Binary64 oracle of the neural radiance cache MLP (vitrum_amd/csrc/nrc.hip).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

Mirrors the device network: the 64-dim input encoding, four 64x64 ReLU
layers and a 3x64 linear output (no biases; input 63 is a constant 1), with
bf16 rounding of weights, inputs, activations and deltas at exactly the
points where the device rounds (MFMA operands) and binary64 everywhere else.
The loss and its output gradient follow
walkaround-hybrid/src/neural/nrc/wgsl/fusedMlp.wgsl.ts:65-120
(nrcRelativeL2OutputGradient); Adam follows
walkaround-hybrid/src/neural/nrc/fusedMlpTrainer.ts:30-84.
`forward_exact`/`backward_exact` are the same network without rounding; a
finite-difference test pins them.
"""
from __future__ import annotations

import numpy as np

N_FEAT, DIMS, WIDTH, OUT, HIDDEN = 16, 64, 64, 3, 4
N_PARAMS = HIDDEN * WIDTH * WIDTH + OUT * WIDTH
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


def encode(features):
    """64-dim encoding of [n, 16] feature records (float32 inputs), binary64."""
    f = np.asarray(features, np.float32)
    n = f.shape[0]
    x = np.zeros((n, DIMS), np.float64)
    for d in range(36):
        t = f[:, d // 12] * np.float32(1 << (d % 12))       # exact in f32
        r = (t - np.float32(2.0) * np.floor(t * np.float32(0.5))).astype(np.float64)
        x[:, d] = np.sin(np.pi * r)
    for d in range(36, 52):
        q, b = (d - 36) >> 2, (d - 36) & 3
        e = (f[:, 3 + q].astype(np.float64) - (b + 0.5) * 0.25) * 4.0
        x[:, d] = np.exp(-0.5 * e * e)
    x[:, 52] = 1.0 - np.exp(-f[:, 7].astype(np.float64))
    x[:, 53:56] = f[:, 8:11]
    x[:, 56:59] = f[:, 11:14]
    x[:, 63] = 1.0
    return x


def forward(w, features):
    """Device-rounded forward. Returns (y [n,3], acts [H0..H4] (bf16 values))."""
    W = [bf16(m) for m in split(w)]
    h = bf16(encode(features))
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
    return loss, np.concatenate([g.ravel() for g in grads])


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


def adam(w, g, m, v, step, lr=1e-2, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update in binary64 (fusedMlpTrainer.ts:30-84)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    w = np.clip(w - lr * mh / (np.sqrt(vh) + eps), -65504.0, 65504.0)
    return w, m, v
