"""CPU restatement of the present pass (test infrastructure, never shipped).

Follows packages/shared-samplers/src/tonemap.ts:14-120 (operators, exposure,
sRGB OETF; binary64 JavaScript numbers) and the present pass
packages/pt-webgpu/src/wgsl/present.wgsl.ts:58-95 (running mean, operator,
OETF unless linear, clamp to [0, 65504], rgba16float store).  Pinned by the
reference's own tonemap.test.ts vectors (tests/test_oracle_present.py).
"""
import numpy as np

TONEMAP_MODE_INDEX = {"aces": 0, "agx": 1, "reinhard": 2, "linear": 3, "none": 4}
F32_MAX = 3.4028234663852886e38


def exposure(c, e):
    """safelyApplyExposure (tonemap.ts:40-48): saturates at the largest f32."""
    c = np.asarray(c, np.float64)
    if e == 0:
        return np.zeros_like(c)
    mag = np.minimum(np.abs(c) * min(e, F32_MAX), F32_MAX)
    return np.where(c == 0, 0.0, np.where(c < 0, -mag, mag))


def aces(x):
    """acesFilmic (tonemap.ts:51-58), Narkowicz 2015."""
    x = np.asarray(x, np.float64)
    v = np.maximum(x, 0.0)
    with np.errstate(over="ignore", invalid="ignore"):
        y = np.clip((v * (2.51 * v + 0.03)) / (v * (2.43 * v + 0.59) + 0.14), 0.0, 1.0)
    return np.where(v <= 0, 0.0, np.where(v > np.sqrt(np.finfo(np.float64).max / 3), 1.0, y))


def reinhard(x):
    """reinhard (tonemap.ts:61-65)."""
    v = np.maximum(np.asarray(x, np.float64), 0.0)
    return np.where(v == 0, 0.0, 1.0 - 1.0 / (1.0 + v))


def agx(x):
    """agx (tonemap.ts:71-79): per-channel log2 sigmoid."""
    v = np.maximum(np.asarray(x, np.float64), 1e-6)
    lx = np.clip((np.log2(v) + 12.47393) / (12.47393 + 4.026069), 0.0, 1.0)
    x2 = lx * lx
    x4 = x2 * x2
    return np.clip(15.5 * x4 * x2 - 40.14 * x4 * lx + 31.96 * x4 - 6.868 * x2 * lx + 0.4298 * x2
                   + 0.1191 * lx - 0.00232, 0.0, 1.0)


OPS = {"aces": aces, "agx": agx, "reinhard": reinhard, "linear": lambda c: np.clip(c, 0.0, 1.0),
       "none": lambda c: np.asarray(c, np.float64)}


def apply_tonemap(rgb, mode, e=1.0):
    """applyTonemap (tonemap.ts:90-110): exposure, then the operator."""
    if mode not in OPS:
        raise ValueError(mode)
    if e < 0:
        raise ValueError("exposure must be non-negative")
    return OPS[mode](exposure(rgb, e))


def linear_to_srgb(c):
    """linearToSrgb (tonemap.ts:115-120), IEC 61966-2-1."""
    v = np.maximum(np.asarray(c, np.float64), 0.0)
    return np.where(v <= 0.0031308, v * 12.92, 1.055 * np.power(v, 1 / 2.4) - 0.055)


def srgb_to_linear(c):
    c = np.asarray(c, np.float64)
    return np.where(c <= 0.04045, c / 12.92, np.power((c + 0.055) / 1.055, 2.4))


def present(mean_rgb, mode="aces", e=1.0, color_space="srgb"):
    """The present pass on a running-mean image (..., 3): tonemap, OETF unless
    linear, clamp to [0, 65504], rounded to binary16 (the rgba16float
    target)."""
    t = apply_tonemap(np.maximum(mean_rgb, 0.0), mode, e)
    if color_space == "srgb":
        t = linear_to_srgb(t)
    return np.clip(t, 0.0, 65504.0).astype(np.float16).astype(np.float64)
