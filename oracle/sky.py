"""Procedural daylight sky, restated for the tests.  TEST INFRASTRUCTURE ONLY:
imported by tests/ as the checker of the product's host-side bake; the
product never imports it.

The Preetham, Shirley and Smits analytic daylight model (SIGGRAPH 1999)
baked into an equirectangular binary32 map, as
packages/shared-samplers/src/preethamSky.ts:161-313 bakes it, with the solar
cap of solarDiscCoverage.ts:48-182 and the Henyey-Greenstein haze lobe of
hgPhase.ts:61-77; the map's environment parameters follow
packages/pt-webgpu/src/scene/environmentPacking.ts:237-282.  Scalar binary64
loops over Python's math module (the platform libm, as the product's host
code calls it), each expression in the reference's operation order.
"""
import math

import numpy as np

SOLAR_RADIUS = 0.00436
HG_LIMIT = 0.999999
EPS = 2.0 ** -52


def _perez(cos_theta, gamma, k):
    c = max(cos_theta, 1e-4)
    cg = math.cos(gamma)
    return (1 + k[0] * math.exp(k[1] / c)) * (1 + k[2] * math.exp(k[3] * gamma) + k[4] * cg * cg)


def coefficients(T):
    """Perez coefficients A..E for luminance Y and chromaticities x, y as
    linear functions of turbidity (Preetham et al. 1999, appendix A.2)."""
    kY = [0.17872 * T - 1.46303, -0.3554 * T + 0.42749, -0.02266 * T + 5.32505,
          0.12064 * T - 2.57705, -0.06696 * T + 0.37027]
    kx = [-0.01925 * T - 0.25922, -0.06651 * T + 0.00081, -0.00041 * T + 0.21247,
          -0.06409 * T - 0.89887, -0.00325 * T + 0.04517]
    ky = [-0.01669 * T - 0.26078, -0.09495 * T + 0.00921, -0.00792 * T + 0.21023,
          -0.04405 * T - 1.65369, -0.01092 * T + 0.05291]
    return kY, kx, ky


def zenith_luminance(T, theta_sun):
    chi = (4.0 / 9.0 - T / 120.0) * (math.pi - 2 * theta_sun)
    return (4.0453 * T - 4.971) * math.tan(chi) - 0.2155 * T + 2.4192


def zenith_chromaticity(T, ts):
    T2, ts2 = T * T, ts * ts
    ts3 = ts2 * ts
    xz = ((0.00166 * ts3 - 0.00375 * ts2 + 0.00209 * ts) * T2 +
          (-0.02903 * ts3 + 0.06377 * ts2 - 0.03202 * ts + 0.00394) * T +
          (0.11693 * ts3 - 0.21196 * ts2 + 0.06052 * ts + 0.25886))
    yz = ((0.00275 * ts3 - 0.0061 * ts2 + 0.00317 * ts) * T2 +
          (-0.04214 * ts3 + 0.0897 * ts2 - 0.04153 * ts + 0.00516) * T +
          (0.15346 * ts3 - 0.26756 * ts2 + 0.0667 * ts + 0.26688))
    return xz, yz


def xyY_to_rgb(x, y, Y):
    """CIE xyY -> linear Rec.709 (sRGB matrix), negatives clamped."""
    if y < 1e-10 or Y < 0:
        return 0.0, 0.0, 0.0
    X = (Y / y) * x
    Z = (Y / y) * (1 - x - y)
    return (max(0.0, 3.2406 * X - 1.5372 * Y - 0.4986 * Z),
            max(0.0, -0.9689 * X + 1.8758 * Y + 0.0415 * Z),
            max(0.0, 0.0557 * X - 0.204 * Y + 1.057 * Z))


def henyey_greenstein(cos_theta, g):
    cos_theta = max(-1.0, min(1.0, cos_theta))
    g = max(-HG_LIMIT, min(HG_LIMIT, g))
    a = abs(g)
    c = cos_theta if g >= 0 else -cos_theta
    om = 1 - a
    den = om * om + 2 * a * (1 - c)
    return (1 / (4 * math.pi)) * (om * (1 + a)) / (den * math.sqrt(den))


def _wrap(phi):
    w = math.fmod(phi, 2 * math.pi)
    return w + 2 * math.pi if w < 0 else w


def solar_coverage(W, H, sun):
    """Solar-cap occupancy per texel against the centre-sampled measure
    sin(theta_c) dtheta dphi (solarDiscCoverage.ts:48-182); (H * W,) float64."""
    ln = math.hypot(*sun)
    sx, sy, sz = sun[0] / ln, sun[1] / ln, sun[2] / ln
    th_sun = math.acos(max(-1.0, min(1.0, sy)))
    sin_sun = math.hypot(sx, sz)
    sun_phi = _wrap(math.atan2(sz, sx) + math.pi)
    cos_r = math.cos(SOLAR_RADIUS)
    cap = 2 * math.pi * (1 - cos_r)
    dth, dph = math.pi / H, 2 * math.pi / W
    cov = [0.0] * (W * H)
    th_min = max(0.0, th_sun - SOLAR_RADIUS)
    th_max = min(math.pi, th_sun + SOLAR_RADIUS)
    row0 = max(0, math.floor(th_min / dth))
    row1 = min(H - 1, math.ceil(th_max / dth) - 1)
    dep = [0.0]
    if sin_sun <= 1e-12:
        north = sy >= 0
        lo_cap, hi_cap = (0.0, SOLAR_RADIUS) if north else (math.pi - SOLAR_RADIUS, math.pi)
        for r in range(row0, row1 + 1):
            lo, hi = max(r * dth, lo_cap), min((r + 1) * dth, hi_cap)
            if hi <= lo:
                continue
            om = dph * (math.cos(lo) - math.cos(hi))
            for x in range(W):
                cov[r * W + x] = om
            dep[0] += om * W
    else:
        def interval(r, a, b, w):
            if b <= a:
                return
            c0 = max(0, math.floor(a / dph))
            c1 = min(W - 1, math.ceil(b / dph) - 1)
            for x in range(c0, c1 + 1):
                ov = max(0.0, min(b, (x + 1) * dph) - max(a, x * dph))
                if ov <= 0:
                    continue
                cov[r * W + x] += w * ov
                dep[0] += w * ov

        def span(r, half, w):
            if half >= math.pi:
                interval(r, 0, 2 * math.pi, w)
                return
            s = _wrap(sun_phi - half)
            e = s + 2 * half
            interval(r, s, min(e, 2 * math.pi), w)
            if e > 2 * math.pi:
                interval(r, 0, e - 2 * math.pi, w)

        target = SOLAR_RADIUS / 64
        for r in range(row0, row1 + 1):
            lo, hi = max(r * dth, th_min), min((r + 1) * dth, th_max)
            if hi <= lo:
                continue
            ext = hi - lo
            steps = max(16, min(512, math.ceil(ext / target)))
            step = ext / steps
            for k in range(steps):
                th = lo + (k + 0.5) * step
                st = math.sin(th)
                w = st * step
                horiz = st * sin_sun
                if horiz <= EPS:
                    if math.cos(th) * sy >= cos_r:
                        span(r, math.pi, w)
                    continue
                t = (cos_r - math.cos(th) * sy) / horiz
                if t <= -1:
                    span(r, math.pi, w)
                elif t < 1:
                    span(r, math.acos(max(-1.0, min(1.0, t))), w)
    if not (dep[0] > 0) or not math.isfinite(dep[0]):
        raise ValueError("the solar cap could not be integrated")
    norm = cap / dep[0]
    for r in range(row0, row1 + 1):
        centre = ((r + 0.5) / H) * math.pi
        measure = dth * dph * max(math.sin(centre), 1e-6)
        for x in range(W):
            cov[r * W + x] = (cov[r * W + x] * norm) / measure
    return np.array(cov)


def _normalise(v):
    ln = math.hypot(*v)
    return (0.0, 1.0, 0.0) if ln < 1e-8 else tuple(c / ln for c in v)


def bake(sun_direction=(0.0, 1.0, 0.0), turbidity=2.0, rayleigh=1.0, mie_coefficient=0.005,
         mie_directional_g=0.8, intensity=1.0, width=256, height=128):
    """(rgb (height, width, 3) float32, luminance integral) of the procedural
    sky, parameters as ProceduralSkyEnvironment names them; the sun direction
    is normalised twice (environmentPacking.ts:245-249, then the bake)."""
    if not (math.isfinite(intensity) and intensity >= 0):
        raise ValueError("intensity")
    intensity = float(np.float32(intensity))
    sun = _normalise(_normalise(tuple(float(c) for c in sun_direction)))
    T = max(1.5, min(30.0, turbidity))
    rayleigh = max(0.0, rayleigh)
    mie_scale = max(0.0, mie_coefficient) * 200
    g = max(-HG_LIMIT, min(HG_LIMIT, mie_directional_g))
    intensity = max(0.0, intensity)
    th_sun = math.acos(max(-1.0, min(1.0, sun[1])))
    kY, kx, ky = coefficients(T)
    zY = max(1e-5, zenith_luminance(T, th_sun)) * rayleigh * intensity
    xz, yz = zenith_chromaticity(T, th_sun)
    nY, nx, ny = _perez(1, th_sun, kY), _perez(1, th_sun, kx), _perez(1, th_sun, ky)
    cov = solar_coverage(width, height, sun)
    out = np.zeros((height * width, 3), np.float32)
    dph = 2 * math.pi / width
    total = 0.0
    for py in range(height):
        theta = ((py + 0.5) / height) * math.pi
        ct, st = math.cos(theta), math.sin(theta)
        omega = dph * (math.cos((py / height) * math.pi) - math.cos(((py + 1) / height) * math.pi))
        for px in range(width):
            phi = ((px + 0.5) / width - 0.5) * (2 * math.pi)
            dx, dy, dz = st * math.cos(phi), ct, st * math.sin(phi)
            cg = max(-1.0, min(1.0, dx * sun[0] + dy * sun[1] + dz * sun[2]))
            gamma = math.acos(cg)
            c = (0.0, 0.0, 0.0)
            if ct >= -0.05:
                sc = max(ct, 1e-4)
                sY = zY * (_perez(sc, gamma, kY) / nY) if nY > 1e-12 else 0.0
                sx = xz * (_perez(sc, gamma, kx) / nx) if nx > 1e-12 else xz
                sy = yz * (_perez(sc, gamma, ky) / ny) if ny > 1e-12 else yz
                fade = min(1.0, (ct + 0.05) / 0.05)
                lum = max(0.0, sY) * fade
                mie_y = mie_scale * henyey_greenstein(cg, g) * 4 * math.pi * zY * fade
                sun_y = cov[py * width + px] * 500 * zY
                Y = lum + mie_y + sun_y
                sky = lum > 1e-12 or Y < 1e-12
                c = xyY_to_rgb(sx if sky else xz, sy if sky else yz, Y)
            i = py * width + px
            out[i] = c
            q = [float(v) for v in out[i]]
            total += max(0.0, (0.2126 * q[0] + 0.7152 * q[1] + 0.0722 * q[2]) * omega)
    return out.reshape(height, width, 3), total
