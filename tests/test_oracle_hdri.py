"""CPU pins of the HDRI environment in the binary64 integrator mirror.

  * importance table (pt-webgpu/src/scene/environmentPacking.ts:283-440): the
    texel pdf integrates to one over the sphere, is proportional to luminance
    to one represented bucket, ignores the intensity (which scales the
    binary32 radiance only), and a black map is no environment;
  * lookup convention (connect.wgsl.ts:126-153, environment.ts:21-56): a
    directly viewed map returns, for every pixel whose footprint lies inside
    one texel, exactly that texel's radiance, at several rotations, the texel
    predicted by an independent numpy restatement of the documented
    convention (+Y at row 0, phi = -pi at column 0, world d looks up RY(-rot) d);
  * estimator: a floor under a map that varies with elevation only, and a
    wall under a map with a bright 'sun' at several rotations, converge to a
    dense hemisphere quadrature of f cos L that shares no code with the NEE
    sampler or its rotation.  A wrong rotation sign on the sampled direction,
    a pdf taken from the wrong texel or a dropped intensity moves the wall's
    mean by far more than the tolerance.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from vitrum_amd import workloads as W

from render_cases import (cpu_frames, directional_albedo_at_normal, env_texel_index, hdri_reflected_radiance,
                          hdri_sun_map, plane_under_sky, sky_view, sky_view_expected, wall_under_hdri)


def _omega(H, W_):
    y = np.repeat(np.arange(H), W_)
    return 2 * np.pi / W_ * (np.cos(y / H * np.pi) - np.cos((y + 1) / H * np.pi))


def test_table_pdf_integrates_to_one_and_follows_luminance():
    tex = hdri_sun_map()
    t = O.environment_table(env_map={"texels": tex, "intensity": 3.0})
    H, W_ = tex.shape[:2]
    om = _omega(H, W_)
    pdf = t["texels"][:, 3].astype(np.float64)
    assert float(np.sum(pdf * om)) == pytest.approx(1.0, abs=1e-6)
    lum = 0.2126 * tex[..., 0] + 0.7152 * tex[..., 1] + 0.0722 * tex[..., 2]
    ideal = (lum.ravel() * om) / float(np.sum(lum.ravel() * om))
    # bucket counts: one per positive texel plus the nearest apportionment of
    # the rest (the f32 cdf holds integer / 2^24 exactly)
    counts = np.diff(t["cdf"].astype(np.float64)) * (1 << 24)
    assert np.array_equal(counts, np.round(counts)) and counts.sum() == 1 << 24
    assert np.abs(counts - (1 + ideal * ((1 << 24) - ideal.size))).max() < 1.0
    np.testing.assert_allclose(pdf, (counts / (1 << 24) / om).astype(np.float32), rtol=0)
    # intensity scales the radiance (one binary32 product), not the pdf
    np.testing.assert_array_equal(t["texels"][:, :3], (tex.reshape(-1, 3) * np.float32(3.0)).astype(np.float32))
    t1 = O.environment_table(env_map={"texels": tex, "intensity": 1.0})
    np.testing.assert_array_equal(t1["texels"][:, 3], t["texels"][:, 3])
    assert t["power"] == pytest.approx(3.0 * t1["power"], rel=1e-6)


def test_black_map_is_no_environment_and_channels_expand():
    assert O.environment_table(env_map={"texels": np.zeros((4, 8, 3), np.float32)}) is None
    grey = np.full((4, 8, 1), 0.5, np.float32)
    t = O.environment_table(env_map={"texels": grey})
    np.testing.assert_array_equal(t["texels"][:, :3], np.full((32, 3), 0.5, np.float32))
    rg = np.zeros((4, 8, 2), np.float32)
    rg[..., 0], rg[..., 1] = 0.25, 0.75
    t = O.environment_table(env_map={"texels": rg})
    assert t["texels"][0, :3].tolist() == [0.25, 0.75, 0.25]


def test_rotation_is_reduced_mod_two_pi_then_rounded():
    tex = hdri_sun_map()
    for rot in (0.0, 1.0, -2.5, 7.0, 100.0):
        t = O.environment_table(env_map={"texels": tex, "rotation_y": rot})
        r = float(np.float32(math.fmod(float(np.float32(rot)), 2 * math.pi)))
        assert t["rot"] == (float(np.float32(math.cos(r))), float(np.float32(math.sin(r))))


def test_lookup_convention_known_directions():
    H, W_ = 16, 32
    r, c = env_texel_index((H, W_), 0.0, [[1, 0, 0], [0, 0, 1], [-1, 0, 1e-9], [0, 1, 0], [0, -1, 0]])
    assert c.tolist()[:3] == [W_ // 2, 3 * W_ // 4, W_ - 1]
    assert r.tolist()[3:] == [0, H - 1]
    # a dome rotated by +pi/2: world +X looks up the map's +Z
    r, c = env_texel_index((H, W_), math.pi / 2, [[1, 0.0, 0]])
    assert c.tolist() == [3 * W_ // 4]


@pytest.mark.parametrize("rot", [0.0, 0.7, -2.0])
def test_directly_viewed_map_is_exact(rot):
    sc = sky_view({"texels": hdri_sun_map(), "intensity": 1.5, "rotation_y": rot})
    res = 24
    bias, _ = W.ray_bias(sc["scene_radius"])
    imgs, _ = cpu_frames(sc, res, 2, float(bias))
    want, inside = sky_view_expected(sc, res)
    assert inside.mean() > 0.5
    for img in imgs:
        np.testing.assert_array_equal(img[inside].astype(np.float32), want[inside])


def test_floor_under_elevation_map_matches_quadrature():
    """A map constant along each row (elevation only): the floor's reflected
    radiance is the row radiances weighted by f cos over each row's band."""
    H, W_ = 16, 32
    rows = np.linspace(2.0, 0.2, H)
    tex = np.repeat(np.stack([rows, rows * 0.8, rows * 0.6], -1)[:, None, :], W_, axis=1).astype(np.float32)
    sc = plane_under_sky(0.5, [0, 0, 0], hdri={"texels": tex, "intensity": 1.0, "rotation_y": 0.4})
    bias, _ = W.ray_bias(sc["scene_radius"])
    imgs, _ = cpu_frames(sc, 32, 32, float(bias))
    mean = imgs.mean(axis=(0, 1, 2))
    lit = lambda th: tex[np.minimum((th / np.pi * H).astype(int), H - 1), 0].astype(np.float64)  # noqa: E731
    expect = directional_albedo_at_normal(0.5, 1.0, 0.0, n_theta=4096, radiance_of_theta=lit)
    np.testing.assert_allclose(mean, expect, rtol=4e-3)


@pytest.mark.parametrize("rot,roughness,metallic", [(math.pi / 2, 1.0, 0.0), (0.0, 1.0, 0.0),
                                                    (-1.0, 0.4, 0.0), (2.2, 0.5, 1.0)])
def test_wall_under_sun_map_matches_quadrature(rot, roughness, metallic):
    hdri = {"texels": hdri_sun_map(), "intensity": 1.0, "rotation_y": rot}
    sc = wall_under_hdri(hdri, 0.5, roughness, metallic)
    bias, _ = W.ray_bias(sc["scene_radius"])
    imgs, _ = cpu_frames(sc, 32, 48, float(bias))
    px = imgs.reshape(imgs.shape[0], -1, 3).mean(axis=1)  # per-frame means
    mean = px.mean(axis=0)
    sem = px.std(axis=0, ddof=1) / math.sqrt(px.shape[0])
    expect = hdri_reflected_radiance(hdri, [1, 0, 0], 0.5, roughness, metallic)
    assert (np.abs(mean - expect) <= 4 * sem + 4e-3 * expect).all(), (mean, expect, sem)


def test_wall_quadrature_sees_the_rotation():
    """The wall test has power: the sun faces the wall only at rotation pi/2."""
    tex = hdri_sun_map()
    lit = hdri_reflected_radiance({"texels": tex, "rotation_y": math.pi / 2}, [1, 0, 0], 0.5, 1.0, 0.0,
                                  n_theta=256, n_phi=512)
    away = hdri_reflected_radiance({"texels": tex, "rotation_y": -math.pi / 2}, [1, 0, 0], 0.5, 1.0, 0.0,
                                   n_theta=256, n_phi=512)
    assert (lit > 3 * away).all()
