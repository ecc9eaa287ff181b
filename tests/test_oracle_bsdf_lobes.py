"""CPU pins of the oracle's extension lobes (clearcoat, sheen, anisotropic
GGX, KHR specular; bsdf.wgsl.ts:392-1489, 2920-2997), through the C hooks
orc_bsdf_eval / orc_bsdf_pdf / orc_bsdf_sample over material_table records.

  * sampling and density agree: for each lobe mix, a histogram of sampled
    directions over (cos theta, phi) bins matches the density integrated over
    the bins (chi-square), and the density's hemisphere integral equals the
    fraction of samples above the horizon -- a wrong sheen exponent, swapped
    anisotropy axes or a wrong lobe probability fails this;
  * the selection frequencies of the four lobes are the joint probabilities
    of brdfRepresentedExtensionLobeProbabilities (cc / (1 + cc + sheen),
    ...);
  * a smooth clearcoat over a black base reflects the coat's F0 = 0.04 times
    clearcoat at normal incidence (Schlick with F0 0.04 at grazing view);
  * anisotropy -> 0 is continuous with the isotropic lobe, a rotation by pi is
    the identity, and a rotation by pi/2 equals rotating the directions by
    -pi/2 about the normal (the frame t' = cos t + sin b);
  * energy: no lobe mix reflects more than it receives (white furnace
    albedo <= 1), and every extension lobe at 0 leaves the base BSDF.
"""
import math

import numpy as np
import pytest
from scipy import stats

from oracle import oracle as O

N = np.array([0.0, 1.0, 0.0])


def row(**m):
    return O.material_table([m])[0]


def unit(theta, phi):
    """A direction at polar angle theta from +y, azimuth phi in the x-z plane."""
    return np.array([math.sin(theta) * math.cos(phi), math.cos(theta), math.sin(theta) * math.sin(phi)])


def hemisphere_grid(nt=192, nphi=384):
    ct = (np.arange(nt) + 0.5) / nt
    ph = (np.arange(nphi) + 0.5) / nphi * 2 * math.pi
    st = np.sqrt(1 - ct * ct)
    dirs = np.stack([np.outer(st, np.cos(ph)), np.outer(ct, np.ones(nphi)), np.outer(st, np.sin(ph))], -1)
    return dirs.reshape(-1, 3), (1.0 / nt) * (2 * math.pi / nphi)


def integrate(fn, nt=192, nphi=384):
    dirs, dw = hemisphere_grid(nt, nphi)
    return sum(fn(d) for d in dirs) * dw


MIXES = {
    "clearcoat": dict(baseColor=[0.6, 0.3, 0.2], roughness=0.7, clearcoat=0.8, clearcoatRoughness=0.2),
    "sheen": dict(baseColor=[0.2, 0.5, 0.3], roughness=0.9, sheen=1.0, sheenRoughness=0.5,
                  sheenColor=[1.0, 0.8, 0.9]),
    "aniso": dict(baseColor=[0.9, 0.8, 0.7], roughness=0.45, metallic=1.0, anisotropy=0.8, anisotropyRotation=0.6),
    "all": dict(baseColor=[0.5, 0.5, 0.6], roughness=0.5, metallic=0.3, clearcoat=0.5, clearcoatRoughness=0.3,
                sheen=0.7, sheenRoughness=0.35, sheenColor=[0.9, 0.9, 0.6], anisotropy=0.5, anisotropyRotation=2.0,
                specularColor=[1.2, 0.9, 1.0], specularIntensity=0.7),
}


@pytest.mark.parametrize("name", list(MIXES))
def test_sampled_directions_follow_the_density(name):
    m = row(**MIXES[name])
    wo = unit(0.7, 0.4)
    count = 60000
    wi, lobe = O.bsdf_sample(m, N.astype(np.float32), wo.astype(np.float32), count, seed=12345)
    wi = wi.astype(np.float64)
    up = wi @ N > 1e-5
    nt, nphi = 12, 24
    ct = wi[up] @ N
    ph = np.mod(np.arctan2(wi[up, 2], wi[up, 0]), 2 * math.pi)
    hist, _, _ = np.histogram2d(ct, ph, bins=[nt, nphi], range=[[0, 1], [0, 2 * math.pi]])
    # expected counts: the density integrated over each bin (4 x 4 midpoints)
    exp = np.zeros((nt, nphi))
    sub = 4
    for i in range(nt):
        for j in range(nphi):
            acc = 0.0
            for a in range(sub):
                for b in range(sub):
                    c = (i + (a + 0.5) / sub) / nt
                    p = (j + (b + 0.5) / sub) / nphi * 2 * math.pi
                    acc += O.bsdf_pdf(m, N, wo, unit(math.acos(c), p))
            exp[i, j] = acc / sub / sub * (1.0 / nt) * (2 * math.pi / nphi) * count
    mass = exp.sum()
    # the density's hemisphere mass is the fraction of samples above the horizon
    assert mass / count == pytest.approx(up.mean(), abs=0.01)
    keep = exp > 30
    chi2 = float((((hist - exp) ** 2) / exp)[keep].sum())
    dof = int(keep.sum()) - 1
    assert stats.chi2.sf(chi2, dof) > 1e-4, f"{name}: chi2 {chi2:.1f} over {dof} bins"


def test_lobe_selection_frequencies():
    cc, sh = 0.6, 0.3
    m = row(baseColor=[0.5, 0.5, 0.5], roughness=0.5, clearcoat=cc, clearcoatRoughness=0.1, sheen=sh,
            sheenRoughness=0.5, sheenColor=[1, 1, 1])
    _, lobe = O.bsdf_sample(m, N.astype(np.float32), unit(0.3, 0.0).astype(np.float32), 40000, seed=7)
    total = 1 + cc + sh
    for k, p in ((2, cc / total), (3, sh / total)):
        f = np.mean(lobe == k)
        assert abs(f - p) < 4 * math.sqrt(p * (1 - p) / lobe.size), (k, f, p)
    assert abs(np.mean(lobe <= 1) - 1 / total) < 0.01
    # both extension scalars above one: the base keeps 1 / (1 + cc + sheen)
    m = row(clearcoat=1.0, sheen=1.0, sheenColor=[1, 1, 1], sheenRoughness=0.5)
    _, lobe = O.bsdf_sample(m, N.astype(np.float32), unit(0.3, 0.0).astype(np.float32), 30000, seed=9)
    assert abs(np.mean(lobe <= 1) - 1 / 3) < 0.012
    # no extension: every sample is a base lobe
    _, lobe = O.bsdf_sample(row(roughness=0.5), N.astype(np.float32), unit(0.3, 0.0).astype(np.float32), 2000)
    assert lobe.max() <= 1


def test_smooth_clearcoat_reflects_the_coat_fresnel():
    """Base black, dielectric F0 0 (specularIntensity 0): the BSDF is the coat
    alone; a near-smooth coat's directional albedo is its Schlick Fresnel with
    F0 0.04 (times clearcoat): 0.04 at normal incidence and
    0.04 + 0.96 (1 - cos)^5 at an oblique view."""
    for cc in (1.0, 0.5):
        m = row(baseColor=[0, 0, 0], roughness=1.0, specularIntensity=0.0, clearcoat=cc, clearcoatRoughness=0.1)
        for theta in (0.0, 1.0):
            wo = unit(theta, 0.3)
            # Monte Carlo with the sampler (its agreement with the density is
            # pinned above): the mean of f cos / pdf
            wi, _ = O.bsdf_sample(m, N.astype(np.float32), wo.astype(np.float32), 20000, seed=11)
            acc = 0.0
            for d in wi.astype(np.float64):
                pdf = O.bsdf_pdf(m, N, wo, d)
                if pdf > 0:
                    acc += O.bsdf_eval(m, N, wo, d)[0] * (d @ N) / pdf
            albedo = acc / wi.shape[0]
            x = 1 - math.cos(theta)
            want = cc * (0.04 + 0.96 * x ** 5)
            assert albedo == pytest.approx(want, rel=0.03), (cc, theta, albedo, want)


def test_anisotropy_limits_and_rotation():
    base = dict(baseColor=[0.8, 0.6, 0.4], roughness=0.4, metallic=0.5)
    iso = row(**base)
    tiny = row(**base, anisotropy=1e-6, anisotropyRotation=0.9)
    rng = np.random.default_rng(3)
    for _ in range(50):
        wo = unit(rng.uniform(0, 1.4), rng.uniform(0, 6.28))
        wi = unit(rng.uniform(0, 1.4), rng.uniform(0, 6.28))
        np.testing.assert_allclose(O.bsdf_eval(tiny, N, wo, wi), O.bsdf_eval(iso, N, wo, wi), rtol=1e-4, atol=1e-9)
        assert O.bsdf_pdf(tiny, N, wo, wi) == pytest.approx(O.bsdf_pdf(iso, N, wo, wi), rel=1e-4, abs=1e-9)
    a0 = row(**base, anisotropy=0.9, anisotropyRotation=0.0)
    api = row(**base, anisotropy=0.9, anisotropyRotation=math.pi)
    ahalf = row(**base, anisotropy=0.9, anisotropyRotation=math.pi / 2)
    # the unrotated frame of buildOnb (bsdf.wgsl.ts:1632-1640); rotating the
    # material by pi/2 rotates its lobe by pi/2 about the normal in that frame
    up = np.array([1.0, 0.0, 0.0]) if abs(N[1]) > 0.999 else np.array([0.0, 1.0, 0.0])
    t0 = np.cross(up, N); t0 /= np.linalg.norm(t0)
    b0 = np.cross(N, t0)
    def about_n(v, a):
        vt, vb, vn = v @ t0, v @ b0, v @ N
        c, s = math.cos(a), math.sin(a)
        return (c * vt - s * vb) * t0 + (s * vt + c * vb) * b0 + vn * N
    differs = 0
    for _ in range(50):
        wo = unit(rng.uniform(0.1, 1.3), rng.uniform(0, 6.28))
        wi = unit(rng.uniform(0.1, 1.3), rng.uniform(0, 6.28))
        f0 = O.bsdf_eval(a0, N, wo, wi)
        np.testing.assert_allclose(O.bsdf_eval(api, N, wo, wi), f0, rtol=1e-6, atol=1e-12)
        # the material rotated by pi/2, seen from directions rotated by pi/2
        np.testing.assert_allclose(O.bsdf_eval(ahalf, N, about_n(wo, math.pi / 2), about_n(wi, math.pi / 2)), f0,
                                   rtol=1e-6, atol=1e-12)
        differs += not np.allclose(O.bsdf_eval(ahalf, N, wo, wi), f0, rtol=1e-3)
    assert differs > 25  # the lobe is anisotropic


@pytest.mark.parametrize("name", list(MIXES))
def test_white_furnace_albedo_is_at_most_one(name):
    """Away from grazing views (where the base BSDF itself -- Lambert with a
    (1 - F) factor plus Kulla-Conty -- returns up to 1 % more than it
    receives), no lobe mix reflects more than it receives."""
    spec = dict(MIXES[name])
    spec["baseColor"] = [1.0, 1.0, 1.0]
    spec.pop("specularColor", None)
    m = row(**spec)
    for theta in (0.2, 0.9):
        wo = unit(theta, 1.1)
        albedo = integrate(lambda d: O.bsdf_eval(m, N, wo, d) * (d @ N), 96, 192)
        assert np.all(albedo <= 1.0 + 1e-3), (name, theta, albedo)
        assert np.all(albedo > 0.3)


@pytest.mark.parametrize("theta", [0.2, 0.9, 1.35])
def test_clearcoat_layering_conserves_the_coat_fresnel(theta):
    """KHR_materials_clearcoat layering: layered = base * (1 - w) + coat with
    w = clearcoat * F_coat(|V.N|); the coat lobe (G <= 1) returns at most w,
    so layered - base * (1 - w) lies in [0, w].  A missing base attenuation or
    an unweighted coat breaks the bound."""
    base_spec = dict(baseColor=[1.0, 1.0, 1.0], roughness=0.6)
    cc = 0.9
    layered, base = row(**base_spec, clearcoat=cc, clearcoatRoughness=0.25), row(**base_spec)
    wo = unit(theta, 0.7)
    x = 1 - math.cos(theta)
    w = cc * (0.04 + 0.96 * x ** 5)
    a_l = integrate(lambda d: O.bsdf_eval(layered, N, wo, d)[0] * (d @ N), 128, 256)
    a_b = integrate(lambda d: O.bsdf_eval(base, N, wo, d)[0] * (d @ N), 128, 256)
    coat = a_l - a_b * (1 - w)
    assert -1e-4 <= coat <= w + 1e-4, (theta, coat, w)
    assert coat > 0.5 * w  # a rough coat still returns most of its Fresnel


def test_zero_extensions_leave_the_base():
    base = dict(baseColor=[0.7, 0.4, 0.3], roughness=0.35, metallic=0.2)
    a, b = row(**base), row(**base, clearcoat=0, sheen=0, sheenColor=[1, 1, 1], anisotropy=0,
                            specularColor=[1, 1, 1], specularIntensity=1)
    rng = np.random.default_rng(5)
    for _ in range(20):
        wo, wi = unit(rng.uniform(0, 1.4), rng.uniform(0, 6.3)), unit(rng.uniform(0, 1.4), rng.uniform(0, 6.3))
        np.testing.assert_array_equal(O.bsdf_eval(b, N, wo, wi), O.bsdf_eval(a, N, wo, wi))
        assert O.bsdf_pdf(b, N, wo, wi) == O.bsdf_pdf(a, N, wo, wi)
    # a sheen layer with no colour still attenuates nothing and adds nothing
    c = row(**base, sheen=1.0, sheenRoughness=0.4)
    wo, wi = unit(0.5, 0.2), unit(0.8, 2.0)
    np.testing.assert_allclose(O.bsdf_eval(c, N, wo, wi), O.bsdf_eval(a, N, wo, wi), rtol=1e-12)
