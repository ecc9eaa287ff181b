"""Procedural sky (ProceduralSkyEnvironment, environment.ts:59-71) through
the product: the host-side Preetham bake (vt_bake_procedural_sky) equals the
restatement bit for bit, and a scene lit by it meets the path-matching parity
contract against the binary64 integrator (the baked map takes the HDRI path;
environmentPacking.ts:222-282)."""
import numpy as np
import pytest

from vitrum_amd import native
from vitrum_amd import workloads as W

from oracle import oracle as O
from render_cases import materials_of, sky_bake
from test_gpu_render import check_single_frame

pytestmark = pytest.mark.gpu

SKIES = [{"sunDirection": [0.4, 0.7, -0.2]},
         {"sunDirection": [0.0, 1.0, 0.0], "mieCoefficient": 0.0},
         {"sunDirection": [1.0, 0.05, 0.0], "turbidity": 7.0, "mieDirectionalG": 0.99995, "intensity": 0.3},
         {"sunDirection": [0.0, -1.0, 0.0], "rayleigh": 2.0},
         {"sunDirection": [0.0, 0.0, 0.0], "turbidity": 100.0, "mieDirectionalG": -3.0}]


@pytest.mark.parametrize("sky", SKIES)
def test_bake_equals_restatement(sky):
    got, integral = native.bake_procedural_sky(sky)
    want = sky_bake(sky)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    from oracle import sky as S
    _, want_integral = S.bake(sun_direction=sky.get("sunDirection", (0, 1, 0)), turbidity=sky.get("turbidity", 2.0),
                              rayleigh=sky.get("rayleigh", 1.0), mie_coefficient=sky.get("mieCoefficient", 0.005),
                              mie_directional_g=sky.get("mieDirectionalG", 0.8), intensity=sky.get("intensity", 1.0))
    assert integral == want_integral


def test_invalid_sky_parameters_are_range_errors(engine):
    for bad in ({"turbidity": float("nan")}, {"intensity": -1.0}, {"sunDirection": [float("inf"), 1, 0]},
                {"mieCoefficient": 1e308}):
        with pytest.raises(native.VtError) as ei:
            native.bake_procedural_sky(bad)
        assert ei.value.status == -2, str(ei.value)
    sc = W.spheres_scene(instances=16, lights=4)
    engine.set_scene(sc)
    info0 = engine.render_info()
    with pytest.raises(native.VtError):
        engine.set_scene(dict(sc, procedural_sky={"rayleigh": float("nan")}))
    assert engine.render_info() == info0


def spheres_under_sky(sky):
    sc = W.spheres_scene(instances=48, lights=4)
    return dict(sc, sky=[0.0, 0.0, 0.0], procedural_sky=sky)


def test_sky_light_tree_equals_oracle(engine):
    sc = spheres_under_sky(SKIES[0])
    engine.set_scene(sc)
    got = engine.light_tree()
    pack = O.pack_scene([dict(p, material_id=i) for i, p in enumerate(sc["primitives"])])
    em = O.build_emitters(pack, O.material_table(materials_of(sc)))
    want = O.scene_light_tree(em, None, env_map={"texels": sky_bake(SKIES[0])})
    assert want is not None and got.shape == want.shape
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("sky", [SKIES[0], SKIES[2]])
def test_sky_paths_match_fp64_integrator(engine, sky):
    gpu, cpu, rate = check_single_frame(engine, spheres_under_sky(sky), 64)
    assert gpu.mean() > 0.0
