"""CPU pins of the binary64 integrator mirror (oracle/integrator.c).

  * represented distributions: the reference's own vectors
    (shared-samplers/src/__tests__/representedDistribution.test.ts:10-55);
  * estimator: a plane under a constant sky converges to L times the BSDF's
    directional albedo, the albedo from a dense quadrature that shares no code
    with any sampling routine (a dropped pdf term, MIS weight or Fresnel factor
    moves the mean by far more than the tolerance);
  * a directly viewed emitter returns its radiance exactly;
  * determinism: the same seed gives the same image and path signatures.
"""
import numpy as np
import pytest

from oracle import oracle as O
from vitrum_amd import workloads as W

from render_cases import SEED, cpu_frames, directional_albedo_at_normal, plane_under_sky


def test_represented_distribution_reference_vectors():
    cdf, counts = O.represented_cdf([2.0 ** -30, 2.0 ** -30, 1.0], counts=True)
    assert counts.tolist() == [1, 1, (1 << 24) - 2]
    assert cdf[0] == 0 and cdf[-1] == 1
    cdf, counts = O.represented_cdf([1, 1, 1], bucket_bits=3, counts=True)
    assert counts.tolist() == [3, 3, 2]
    assert cdf.tolist() == [0, 3 / 8, 6 / 8, 1]
    cdf, counts = O.represented_cdf([0, 0, 0], bucket_bits=3, counts=True)
    assert counts.tolist() == [0, 0, 0] and cdf.tolist() == [0, 0, 0, 0]


def test_represented_distribution_apportionment_is_nearest():
    rng = np.random.default_rng(5)
    w = rng.random(2048) ** 4
    cdf, counts = O.represented_cdf(w, counts=True)
    assert counts.sum() == 1 << 24 and (counts >= 1).all()
    ideal = 1 + w / w.sum() * ((1 << 24) - w.size)
    assert np.abs(counts - ideal).max() < 1.0
    assert (np.diff(cdf) >= 0).all()


def test_constant_env_pdf_integrates_to_one():
    cdf, pdf = O.build_constant_env([1.0, 0.8, 0.6], 64, 32)
    y = np.repeat(np.arange(32), 64)
    omega = 2 * np.pi / 64 * (np.cos(y / 32 * np.pi) - np.cos((y + 1) / 32 * np.pi))
    assert float(np.sum(pdf.astype(np.float64) * omega)) == pytest.approx(1.0, abs=1e-6)
    # constant radiance: pdf is uniform over the sphere to one bucket per texel
    quantum = 1.0 / ((1 << 24) * omega)
    assert (np.abs(pdf - 1 / (4 * np.pi)) <= quantum * 1.0001).all()


@pytest.mark.parametrize("roughness,metallic", [(1.0, 0.0), (0.3, 0.0), (0.5, 1.0)])
def test_plane_under_constant_sky_matches_quadrature(roughness, metallic):
    L = np.array([1.0, 0.8, 0.6])
    sc = plane_under_sky(0.5, L.tolist(), roughness, metallic)
    bias, _ = W.ray_bias(sc["scene_radius"])
    imgs, _ = cpu_frames(sc, 32, 32, float(bias))
    mean = imgs.mean(axis=(0, 1, 2))
    expect = L * directional_albedo_at_normal(0.5, roughness, metallic)
    np.testing.assert_allclose(mean, expect, rtol=3e-3)


def test_quadrature_energy_bounds():
    """The quadrature itself: with a black base only the 4% dielectric specular
    and multiscatter lobes reflect (small, positive albedo); with a white base
    the albedo stays within the energy bound of the diffuse + specular sum."""
    dark = directional_albedo_at_normal(0.0, 1.0, 0.0)
    white = directional_albedo_at_normal(1.0, 1.0, 0.0)
    assert 0.0 < dark < 0.1
    assert 0.9 < white < 1.04


def emitter_card():
    pos = np.array([-5, -5, 0, 5, -5, 0, 5, 5, 0, -5, 5, 0], np.float32)
    nrm = np.tile([0, 0, 1], 4).astype(np.float32)
    idx = np.array([0, 1, 2, 0, 2, 3], np.uint32)
    prim = {"kind": "mesh", "positions": pos, "normals": nrm, "indices": idx,
            "material": {"baseColor": [0, 0, 0], "roughness": 1, "metallic": 0,
                         "emissive": [1.0, 0.5, 0.25], "emissiveIntensity": 2.0}}
    cam = W.look_at_camera(eye=(0, 0, 3), target=(0, 0, 0), up=(0, 1, 0), vfov_deg=30)
    return {"primitives": [prim], "camera": cam, "scene_radius": 8.0, "sky": [0, 0, 0]}


def test_directly_viewed_emitter_is_exact():
    sc = emitter_card()
    imgs, _ = cpu_frames(sc, 16, 2, float(W.ray_bias(8.0)[0]), max_bounces=1)
    np.testing.assert_allclose(imgs.reshape(-1, 3), np.tile([2.0, 1.0, 0.5], (512, 1)), rtol=1e-12)


def test_integrator_is_deterministic():
    sc = W.cornell_scene()
    bias = float(W.ray_bias(sc["scene_radius"])[0])
    a, sa = cpu_frames(sc, 24, 1, bias)
    b, sb = cpu_frames(sc, 24, 1, bias)
    assert np.array_equal(a, b) and np.array_equal(sa, sb)
    c, sc_ = cpu_frames(sc, 24, 2, bias)
    assert not np.array_equal(c[0], c[1])  # frame index reseeds the stream
    assert c[0].mean() > 0.05


def test_light_tree_and_uniform_selection_share_the_expectation():
    """The tree only changes which light NEE picks and the 1/pdf weight
    (selection is compensated outside the MIS, kernel.wgsl.ts:1459-1467), so
    tree and uniform selection estimate the same radiance: per-image means
    over independent frames agree within 4 standard errors, while the tree
    changes the per-path decisions."""
    sc = W.spheres_scene(instances=16, lights=4)
    bias, _ = W.ray_bias(sc["scene_radius"])
    res, frames = 24, 24
    t_img, t_sig = cpu_frames(sc, res, frames, float(bias), max_bounces=4, light_tree=True)
    u_img, u_sig = cpu_frames(sc, res, frames, float(bias), max_bounces=4, light_tree=False)
    assert np.mean(t_sig != u_sig) > 0.05
    tm = t_img.reshape(frames, -1).mean(axis=1)
    um = u_img.reshape(frames, -1).mean(axis=1)
    se = np.sqrt(tm.var(ddof=1) / frames + um.var(ddof=1) / frames)
    assert abs(tm.mean() - um.mean()) <= 4 * se, (tm.mean(), um.mean(), se)


# ---- mesh-area lights: closed forms ---------------------------------------

def _mean_and_bound(imgs, expect, rel=2e-3):
    """Image mean and its allowed deviation: 4 standard errors of the per-path
    values (independent pixels and frames) plus `rel` for the quadrature."""
    v = imgs.reshape(-1, 3)
    se = v.std(axis=0, ddof=1) / np.sqrt(v.shape[0])
    return v.mean(axis=0), 4 * se + rel * np.abs(expect)


def test_polygon_quadrature_matches_lambert_formula():
    """The area quadrature that the disc test integrates with reproduces
    Lambert's closed-form irradiance of a polygon (on axis, and off axis with
    a tilted receiver), and on axis tends to pi sin^2(alpha) for the disc."""
    from render_cases import disc_polygon, polygon_irradiance_lambert, polygon_reflected_radiance
    ring = disc_polygon((0, 1, 0), 0.25, 48)
    tris = [(np.array([0, 1.0, 0]), ring[(i + 1) % 48], ring[i]) for i in range(48)]
    for x, n in (((0, 0, 0), (0, 1, 0)), ((0.4, 0, -0.3), (0.2, 0.9, 0.1))):
        x, n = np.asarray(x, float), np.asarray(n, float) / np.linalg.norm(n)
        q = polygon_reflected_radiance(x, n, None, tris, [1.0])[0] * np.pi
        assert q == pytest.approx(polygon_irradiance_lambert(x, n, ring), rel=2e-5)
    fine = disc_polygon((0, 1, 0), 0.25, 4096)
    e = polygon_irradiance_lambert(np.zeros(3), np.array([0, 1.0, 0]), fine)
    assert e == pytest.approx(np.pi * 0.25 ** 2 / (0.25 ** 2 + 1.0), rel=1e-6)


# (roughness, metallic, light tree, viewed wall emits).  Uniform selection
# keeps the variance small.  The reference's light tree rates a node by the
# distance to its bounding box, and the boxes of merged walls enclose the
# whole interior, so inside an enclosure it picks walls near the shading point
# almost always and the far wall about once in 10^4 draws (with the
# matching 1/pdf): unbiased, but 30x noisier, and its bound is loose.
# The last column is the wall tessellation: with one cell per wall the
# triangles subtend large solid angles, their area pdfs are small, and a
# glossy metal lobe gives the BSDF side of the MIS most of the weight.
ENCLOSURE_CASES = [(1.0, 0.0, False, True, 2), (0.3, 0.0, False, True, 2), (0.5, 1.0, False, True, 2),
                   (0.2, 1.0, False, True, 1), (1.0, 0.0, False, False, 2), (1.0, 0.0, True, False, 2),
                   (0.3, 0.0, True, True, 2)]


@pytest.mark.parametrize("roughness,metallic,light_tree,viewed_wall_emits,n_sub", ENCLOSURE_CASES)
def test_emissive_enclosure_one_bounce(roughness, metallic, light_tree, viewed_wall_emits, n_sub):
    """Inside a closed box of single-sided emitters (radiance Le, albedo a),
    every direction sees Le, so with one bounce the pixel is Le (1 + rho(wo))
    exactly (Le rho(wo) when the viewed wall does not emit), rho from the BSDF
    quadrature.  The estimate runs through light selection (tree or uniform),
    sqrt(r1) triangle sampling, the area to solid angle pdf with the
    light-side cosine, and the power-heuristic MIS of NEE against BSDF samples
    that hit an emitter (the glossy cases give the BSDF side most of the
    weight near the lobe): a wrong area pdf, a missing cosine or a dropped
    MIS weight moves the mean by far more than the bound."""
    from render_cases import emissive_box
    le, a = np.array([1.0, 0.7, 0.4]), 0.6
    sc = emissive_box(le.tolist(), a, roughness, metallic, n_sub=n_sub, viewed_wall_emits=viewed_wall_emits)
    bias, _ = W.ray_bias(sc["scene_radius"])
    imgs, _ = cpu_frames(sc, 24, 16, float(bias), max_bounces=1, light_tree=light_tree)
    expect = le * (float(viewed_wall_emits) + directional_albedo_at_normal(a, roughness, metallic))
    mean, bound = _mean_and_bound(imgs, expect)
    assert (np.abs(mean - expect) <= bound).all(), (mean, expect, bound)


def test_emissive_enclosure_multi_bounce_within_albedo_bounds():
    """With k bounces the pixel lies between Le sum_i rho_min^i and
    Le sum_i rho_max^i (i = 0..k), rho_min/max the extremes of the directional
    albedo over view angles; the k = 3 mean must also exceed the k = 2 mean
    by at least Le rho_min^3."""
    from render_cases import base_bsdf, emissive_box
    le, a = np.array([1.0, 0.7, 0.4]), 0.6
    n = np.array([0.0, 0.0, 1.0])
    rhos = []
    for th in np.linspace(0.0, 1.55, 32):
        wo = np.array([np.sin(th), 0.0, np.cos(th)])
        t = (np.arange(200) + 0.5) / 200 * (np.pi / 2)
        p = (np.arange(400) + 0.5) / 400 * (2 * np.pi)
        T, P = np.meshgrid(t, p, indexing="ij")
        wi = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], axis=-1)
        f = base_bsdf(a, 1.0, 0.0, n, wo, wi)
        rhos.append(float(np.sum(f * np.cos(T) * np.sin(T)) * (np.pi / 2 / 200) * (2 * np.pi / 400)))
    lo, hi = min(rhos), max(rhos)
    sc = emissive_box(le.tolist(), a)
    bias, _ = W.ray_bias(sc["scene_radius"])
    means = {}
    for k in (2, 3):
        imgs, _ = cpu_frames(sc, 24, 16, float(bias), max_bounces=k, light_tree=False)
        mean, bound = _mean_and_bound(imgs, le, rel=0.0)
        means[k] = mean
        assert (mean >= le * sum(lo ** i for i in range(k + 1)) - bound).all(), (k, mean)
        assert (mean <= le * sum(hi ** i for i in range(k + 1)) + bound).all(), (k, mean)
    assert (means[3] - means[2] >= le * lo ** 3 * 0.9).all()


@pytest.mark.parametrize("roughness,metallic,light_tree", [
    (1.0, 0.0, True), (1.0, 0.0, False), (0.4, 0.0, True)])
def test_disc_light_over_plane_matches_quadrature(roughness, metallic, light_tree):
    """A small emitting disc (48-gon) 1 unit above a floor, the floor seen
    30 degrees off its normal: one bounce of NEE + BSDF-hit MIS converges to
    the area quadrature of f Le cos cos / d^2 over the disc's triangles (the
    quadrature is pinned to Lambert's polygon formula)."""
    from render_cases import disc_expected_radiance, disc_over_plane
    le, a = np.array([30.0, 20.0, 10.0]), 0.5
    sc = disc_over_plane(le.tolist(), a, roughness, metallic)
    bias, _ = W.ray_bias(sc["scene_radius"])
    imgs, _ = cpu_frames(sc, 24, 24, float(bias), max_bounces=1, light_tree=light_tree)
    expect = disc_expected_radiance(sc, le, a, roughness, metallic)
    mean, bound = _mean_and_bound(imgs, expect)
    assert (np.abs(mean - expect) <= bound).all(), (mean, expect, bound)
