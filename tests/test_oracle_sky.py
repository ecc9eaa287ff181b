"""CPU pins of the procedural-sky restatement (oracle/sky.py).

  * the reference's own properties of the bake
    (shared-samplers/src/__tests__/preethamSky.test.ts:38-237): finite map,
    normalised represented CDF and per-steradian pdf, black at rayleigh 0,
    intensity scales the integral, the brightest texel lies at the authored
    sun (within a texel diagonal, at two resolutions), the sub-texel sun
    brightens the first row, energy is resolution independent, Mie g beyond
    0.9999 still matters;
  * mathematics: the solar cap deposits exactly its analytic solid angle
    2 pi (1 - cos r) against the bake's centre-sampled measure (polar,
    horizon and oblique suns); the Henyey-Greenstein density integrates to
    one and is 1 / 4 pi at g = 0; the Rec.709 matrix maps the D65 white
    point to equal rgb; the luminance integral is the importance table's
    power (the bake feeds the HDRI path with intensity 1).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from oracle import sky as S


def _omega(H, W):
    y = np.repeat(np.arange(H), W)
    return 2 * np.pi / W * (np.cos(y / H * np.pi) - np.cos((y + 1) / H * np.pi))


def _brightest_direction(rgb):
    H, W = rgb.shape[:2]
    lum = 0.2126 * rgb[..., 0] + 0.7152 * rgb[..., 1] + 0.0722 * rgb[..., 2]
    y, x = np.unravel_index(int(np.argmax(lum)), lum.shape)
    th = (y + 0.5) / H * np.pi
    ph = ((x + 0.5) / W - 0.5) * 2 * np.pi
    return np.array([np.sin(th) * np.cos(ph), np.cos(th), np.sin(th) * np.sin(ph)])


def test_default_bake_is_finite_with_normalised_table():
    rgb, integral = S.bake()
    assert rgb.shape == (128, 256, 3) and np.isfinite(rgb).all() and (rgb >= 0).all()
    assert integral > 0
    t = O.environment_table(env_map={"texels": rgb})
    assert t["cdf"][0] == 0 and t["cdf"][-1] == 1
    # the importance power is the bake's luminance integral
    assert t["power"] == integral


def test_pdf_is_the_represented_pmf_over_solid_angle():
    rgb, _ = S.bake(width=32, height=16)
    t = O.environment_table(env_map={"texels": rgb})
    om = _omega(16, 32)
    pmf = t["cdf"][1:].astype(np.float64) - t["cdf"][:-1].astype(np.float64)
    np.testing.assert_array_equal(t["texels"][:, 3], (pmf / om).astype(np.float32))
    assert float(np.sum(t["texels"][:, 3].astype(np.float64) * om)) == pytest.approx(1.0, abs=1e-6)


def test_rayleigh_zero_is_black_and_intensity_scales():
    rgb, integral = S.bake(rayleigh=0.0, width=32, height=16)
    assert integral == 0 and not rgb.any()
    assert O.environment_table(env_map={"texels": rgb}) is None
    _, unit = S.bake(width=32, height=16)
    _, scaled = S.bake(intensity=3.5, width=32, height=16)
    assert scaled / unit == pytest.approx(3.5, abs=5e-6)


def test_horizon_sun_is_the_brightest_texel():
    rgb, _ = S.bake(sun_direction=(1.0, 0.05, 0.0), width=64, height=32)
    d = _brightest_direction(rgb)
    th, ph = math.acos(d[1]), math.atan2(d[2], d[0])
    assert abs(th - math.pi / 2) < 0.3 and abs(ph) < 0.3


def test_subtexel_polar_sun_brightens_the_first_row():
    rgb, _ = S.bake(sun_direction=(0.0, 1.0, 0.0), mie_coefficient=0.0)
    lum = 0.2126 * rgb[..., 0] + 0.7152 * rgb[..., 1] + 0.0722 * rgb[..., 2]
    assert lum[0].mean() > 2 * lum[1].mean()


@pytest.mark.parametrize("sun", [(0, 1, 0), (1, 0.05, 0), (0.4, 0.7, -0.2)])
def test_energy_and_sun_position_hold_across_resolutions(sun):
    s = np.asarray(sun, np.float64) / np.linalg.norm(sun)
    lo_rgb, lo = S.bake(sun_direction=tuple(s), width=64, height=32)
    hi_rgb, hi = S.bake(sun_direction=tuple(s), width=256, height=128)
    assert 0.94 < hi / lo < 1.06
    for rgb in (lo_rgb, hi_rgb):
        H, W = rgb.shape[:2]
        err = math.acos(max(-1.0, min(1.0, float(_brightest_direction(rgb) @ s))))
        assert err < math.hypot(math.pi / H, 2 * math.pi / W)


def test_mie_asymmetry_beyond_the_old_cutoff_matters():
    a, ia = S.bake(sun_direction=(1.0, 0.05, 0.0), mie_directional_g=0.9999, width=32, height=16)
    b, ib = S.bake(sun_direction=(1.0, 0.05, 0.0), mie_directional_g=0.99995, width=32, height=16)
    assert ia != ib and not np.array_equal(a, b)


@pytest.mark.parametrize("sun,W,H", [((0, 1, 0), 256, 128), ((0, -1, 0), 64, 32), ((1, 0, 0), 256, 128),
                                     ((0.4, 0.7, -0.2), 512, 256), ((-0.3, 0.2, 0.9), 16, 8)])
def test_solar_cap_deposits_its_solid_angle(sun, W, H):
    cov = S.solar_coverage(W, H, sun).reshape(H, W)
    centre = (np.arange(H) + 0.5) / H * np.pi
    measure = (np.pi / H) * (2 * np.pi / W) * np.maximum(np.sin(centre), 1e-6)
    total = float(np.sum(cov * measure[:, None]))
    assert total == pytest.approx(2 * np.pi * (1 - math.cos(S.SOLAR_RADIUS)), rel=1e-12)
    # deposited only within the cap's texels: around the sun's row
    rows = np.nonzero(cov.any(axis=1))[0]
    th = math.acos(sun[1] / math.hypot(*sun))
    assert abs((rows.mean() + 0.5) / H * np.pi - th) < np.pi / H + S.SOLAR_RADIUS


@pytest.mark.parametrize("g", [0.0, 0.5, -0.7, 0.95])
def test_henyey_greenstein_is_a_density(g):
    n = 200000
    mu = -1 + (np.arange(n) + 0.5) * (2.0 / n)
    vals = np.array([S.henyey_greenstein(float(m), g) for m in mu[::50]])
    # the closed-form density (1 - g^2) / (4 pi (1 + g^2 - 2 g mu)^1.5)
    want = (1 - g * g) / (4 * np.pi * (1 + g * g - 2 * g * mu[::50]) ** 1.5)
    np.testing.assert_allclose(vals, want, rtol=1e-10)
    full = (1 - g * g) / (4 * np.pi * (1 + g * g - 2 * g * mu) ** 1.5)
    assert float(np.sum(full) * (2.0 / n) * 2 * np.pi) == pytest.approx(1.0, rel=1e-4)
    if g == 0.0:
        assert S.henyey_greenstein(0.3, 0.0) == pytest.approx(1 / (4 * np.pi), rel=1e-15)


def test_rec709_matrix_maps_d65_to_white():
    r, g, b = S.xyY_to_rgb(0.3127, 0.3290, 1.0)
    assert r == pytest.approx(1.0, abs=2e-3) and g == pytest.approx(1.0, abs=2e-3) and b == pytest.approx(1.0, abs=2e-3)
    assert S.xyY_to_rgb(0.3, 0.0, 1.0) == (0.0, 0.0, 0.0)
