"""HDRI environments on the device (HdriEnvironment, environment.ts:21-56;
environmentPacking.ts:283-440; connect.wgsl.ts:126-219).

  * a directly viewed map: every pixel whose footprint lies inside one texel
    holds exactly that texel's binary32 radiance x intensity, at several
    rotations (texel predicted by the documented convention, in numpy);
  * the same path-matching parity contract as the constant sky against the
    binary64 integrator, with a rotated, intensity-scaled map beside mesh and
    analytic lights, with the light tree and with uniform selection;
  * a wall under a map with a bright sun converges to a dense hemisphere
    quadrature of f cos L (independent of every sampling routine);
  * the packed light tree's environment leaf equals the oracle's;
  * invalid maps are RangeErrors that keep the previous scene; a map that
    is black everywhere renders as no environment.
"""
import math

import numpy as np
import pytest

from vitrum_amd import native
from vitrum_amd import workloads as W

from oracle import oracle as O
from render_cases import (cpu_frames, hdri_reflected_radiance, hdri_sun_map, materials_of, sky_view,
                          sky_view_expected, wall_under_hdri)
from test_gpu_render import check_single_frame, gpu_frames

pytestmark = pytest.mark.gpu


def spheres_under_hdri(rot=0.9, intensity=1.3, emitters=True):
    sc = W.spheres_scene(instances=48, lights=4)
    sc = dict(sc, sky=[0.0, 0.0, 0.0], hdri={"texels": hdri_sun_map(), "intensity": intensity, "rotation_y": rot})
    if emitters:
        sc["emitters"] = [{"kind": "point", "position": [0.0, 6.0, 2.0], "color": [1.0, 0.9, 0.8], "intensity": 20.0},
                          {"kind": "directional", "direction": [-0.3, -1.0, -0.2], "color": [1, 1, 1],
                           "intensity": 0.5}]
    return sc


@pytest.mark.parametrize("rot", [0.0, 0.7, -2.0])
def test_directly_viewed_map_is_exact(engine, rot):
    sc = sky_view({"texels": hdri_sun_map(), "intensity": 1.5, "rotation_y": rot})
    res = 48
    img, _ = gpu_frames(engine, sc, res, 4)
    want, inside = sky_view_expected(sc, res)
    assert inside.mean() > 0.5
    np.testing.assert_array_equal(img[inside].astype(np.float32), want[inside])


def test_hdri_paths_match_fp64_integrator(engine):
    sc = spheres_under_hdri()
    gpu, cpu, rate = check_single_frame(engine, sc, 64)
    assert gpu.mean() > 0.01


def test_hdri_only_paths_match_fp64_integrator(engine):
    sc = spheres_under_hdri(rot=-2.4, intensity=0.7, emitters=False)
    check_single_frame(engine, sc, 64)


@pytest.fixture(scope="module")
def uniform_engine():
    eng = native.Engine(0, light_selection=1)
    yield eng
    eng.close()


def test_hdri_uniform_selection_paths_match(uniform_engine):
    check_single_frame(uniform_engine, spheres_under_hdri(), 48, light_tree=False)


@pytest.mark.parametrize("rot,roughness,metallic", [(math.pi / 2, 1.0, 0.0), (2.2, 0.5, 1.0)])
def test_wall_under_sun_map_converges_to_quadrature(engine, rot, roughness, metallic):
    hdri = {"texels": hdri_sun_map(), "intensity": 1.0, "rotation_y": rot}
    sc = wall_under_hdri(hdri, 0.5, roughness, metallic)
    res, frames = 32, 256
    img, _ = gpu_frames(engine, sc, res, frames)
    mean = img.reshape(-1, 3).mean(axis=0)
    expect = hdri_reflected_radiance(hdri, [1, 0, 0], 0.5, roughness, metallic)
    np.testing.assert_allclose(mean, expect, rtol=1e-2)


def test_light_tree_environment_leaf_equals_oracle(engine):
    sc = spheres_under_hdri()
    engine.set_scene(sc)
    got = engine.light_tree()
    pack = O.pack_scene([dict(p, material_id=i) for i, p in enumerate(sc["primitives"])])
    em = O.build_emitters(pack, O.material_table(materials_of(sc)))
    want = O.scene_light_tree(em, None, analytic=O.pack_analytic(sc["emitters"]), env_map=sc["hdri"])
    assert want is not None and got.shape == want.shape
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_invalid_maps_are_range_errors_and_keep_the_scene(engine):
    good = spheres_under_hdri()
    engine.set_scene(good)
    info0 = engine.render_info()
    tex = hdri_sun_map()
    neg = tex.copy()
    neg[3, 4, 1] = -1.0
    nan = tex.copy()
    nan[0, 0, 0] = np.nan
    for hdri in ({"texels": neg}, {"texels": nan}, {"texels": tex, "intensity": -1.0},
                 {"texels": tex, "intensity": float("inf")}, {"texels": tex, "rotation_y": float("nan")},
                 {"texels": tex, "intensity": 1e38}, {"texels": np.zeros((2, 4, 5), np.float32) + 1}):
        with pytest.raises(native.VtError) as ei:
            engine.set_scene(dict(good, hdri=hdri))
        assert ei.value.status == -2, str(ei.value)
        assert engine.render_info() == info0


def test_black_map_renders_as_no_environment(engine):
    base = spheres_under_hdri()
    res = 32
    black = dict(base, hdri={"texels": np.zeros((8, 16, 3), np.float32)})
    none = dict(base)
    none.pop("hdri")
    a, _ = gpu_frames(engine, black, res, 2)
    b, _ = gpu_frames(engine, none, res, 2)
    assert np.array_equal(a, b)
