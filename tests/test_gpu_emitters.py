"""Analytic emitters on the device (directional, point, spot, rect and disc
areas): per-path agreement with the binary64 integrator under the same seeds
(path signatures, radiance within 1e-4), the closed forms of the oracle's
own pins, and the packed light tree over mixed light kinds bit for bit."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from vitrum_amd import workloads as W

from render_cases import (SEED, cpu_frames, floor_with_emitters, footprint_mean, point_light_radiance,
                          rect_light_radiance)

pytestmark = pytest.mark.gpu

MIXED = [{"kind": "directional", "color": [1, 1, 1], "intensity": 0.5, "direction": [0.3, -1, 0.2],
          "angularDiameter": 0.1},
         {"kind": "point", "color": [1, 0.5, 0.2], "intensity": 20, "position": [1, 2, 0]},
         {"kind": "spot", "color": [0.2, 0.5, 1], "intensity": 30, "position": [-1, 2, 0.5],
          "direction": [0, -1, 0], "angle": 0.8, "penumbra": 0.3},
         {"kind": "rect-area", "color": [1, 1, 1], "intensity": 5, "position": [0, 1.5, 1],
          "uAxis": [0.3, 0, 0], "vAxis": [0, 0, 0.3]},
         {"kind": "disc-area", "color": [1, 1, 0.5], "intensity": 5, "position": [0.5, 1.2, -1],
          "normal": [0, -1, 0], "radius": 0.2},
         {"kind": "point", "color": [1, 1, 1], "intensity": 2, "position": [0, 0.6, 0], "castShadow": False}]


def _scene():
    """The Cornell box plus every analytic kind and a dim sky."""
    sc = W.cornell_scene(4096)
    sc = dict(sc, emitters=[dict(e, position=[e["position"][0] * 0.5, 1.0 + 0.4 * e["position"][1],
                                              e["position"][2] * 0.5]) if "position" in e else e for e in MIXED])
    sc["sky"] = [0.05, 0.05, 0.06]
    return sc


def _gpu(engine, sc, res, frames, bounces=8):
    engine.set_scene(sc)
    engine.reset()
    ivp = W.inv_view_proj(sc["camera"], res, res)
    for f in range(frames):
        engine.render_frame(ivp, res, res, f, SEED, bounces)
    return engine.capture(res, res)[..., :3].astype(np.float64), engine.capture_signatures(res, res)


@pytest.mark.parametrize("uniform", [False, True])
def test_mixed_emitters_paths_match_integrator(engine, uniform):
    from vitrum_amd import native
    eng = native.Engine(0, light_selection=1) if uniform else engine
    try:
        sc = _scene()
        res = 48
        gpu, gsig = _gpu(eng, sc, res, 1)
        bias = eng.render_info()["ray_bias"]
        cpu, csig = cpu_frames(sc, res, 1, bias, light_tree=not uniform)
        cpu, csig = cpu[0], csig[0]
        match = gsig == csig
        assert match.mean() >= 0.999, match.mean()
        rel = (np.abs(gpu - cpu) / (np.abs(cpu) + 1e-6)).max(axis=2)[match]
        print(f"path match {match.mean():.5f}; within 1e-4: {np.mean(rel <= 1e-4):.5f}; max {rel.max():.3g}",
              flush=True)
        ok = (np.abs(gpu - cpu) <= 1e-4 * np.abs(cpu) + 1e-7).all(axis=2) | ~match
        assert ok.mean() >= 0.999, ok.mean()
        assert gpu.mean() > 0.01
    finally:
        if uniform:
            eng.close()


def test_mixed_emitters_light_tree_bytes_equal_oracle(engine):
    from render_cases import materials_of
    sc = _scene()
    engine.set_scene(sc)
    got = engine.light_tree()
    pack = O.pack_scene([dict(p, material_id=i) for i, p in enumerate(sc["primitives"])])
    em = O.build_emitters(pack, O.material_table(materials_of(sc)))
    want = O.scene_light_tree(em, sc["sky"], analytic=O.pack_analytic(sc["emitters"]))
    assert got.shape == want.shape
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def _mean(engine, sc, frames, res=32):
    img, _ = _gpu(engine, sc, res, frames, bounces=1)
    v = img.reshape(-1, 3)
    return v.mean(axis=0), 4 * v.std(axis=0, ddof=1) / np.sqrt(v.shape[0])


def test_point_and_spot_closed_forms(engine):
    bsdf = (0.5, 1.0, 0.0)
    L = {"kind": "point", "color": [1.0, 0.8, 0.6], "intensity": 30.0, "position": [0.3, 2.0, 0.2]}
    sc = floor_with_emitters([L], view_deg=25.0)
    mean, _ = _mean(engine, sc, 1)
    np.testing.assert_allclose(mean, footprint_mean(sc, lambda x, wo: point_light_radiance(x, wo, L, bsdf)),
                               rtol=3e-4)
    S = {"kind": "spot", "color": [1.0, 1.0, 1.0], "intensity": 40.0, "position": [0.0, 2.0, 0.0],
         "direction": [0.4, -1.0, 0.0], "angle": 0.5, "penumbra": 0.4}
    sc = floor_with_emitters([S])
    mean, _ = _mean(engine, sc, 1)
    np.testing.assert_allclose(mean, footprint_mean(sc, lambda x, wo: point_light_radiance(x, wo, S, bsdf)),
                               rtol=3e-4)


@pytest.mark.parametrize("kind", ["rect-area", "disc-area"])
def test_area_light_closed_forms(engine, kind):
    bsdf = (0.9, 1.0, 0.0)
    if kind == "rect-area":
        L = {"kind": kind, "color": [1.0, 1.0, 1.0], "intensity": 3.0, "position": [0.0, 1.0, -0.8],
             "uAxis": [1.5, 0.0, 0.0], "vAxis": [0.0, 0.0, 1.5]}
    else:
        L = {"kind": kind, "color": [1.0, 1.0, 1.0], "intensity": 3.0, "position": [0.0, 1.0, -0.8],
             "normal": [0.0, -1.0, 0.0], "radius": 1.6}
    sc = floor_with_emitters([L], albedo=0.9, roughness=1.0, metallic=0.0, view_deg=30.0)
    mean, se = _mean(engine, sc, 64)
    want = footprint_mean(sc, lambda x, wo: rect_light_radiance(x, wo, L, bsdf, m=96), res=4)
    assert (np.abs(mean - want) <= se + 3e-3 * want).all(), (mean, want, se)


def test_directional_light_closed_form(engine):
    L = {"kind": "directional", "color": [1.0, 0.9, 0.7], "intensity": 2.0, "direction": [0.2, -1.0, 0.1]}
    sc = floor_with_emitters([L], view_deg=20.0)
    mean, _ = _mean(engine, sc, 1)
    from render_cases import base_bsdf
    wi = -np.asarray(L["direction"]) / np.linalg.norm(L["direction"])
    irr = (np.asarray(L["color"], np.float32) * np.float32(2.0)).astype(np.float64)
    want = footprint_mean(sc, lambda x, wo: base_bsdf(0.5, 1.0, 0.0, np.array([0, 1.0, 0]), wo, wi[None, :])[0]
                          * irr * wi[1])
    np.testing.assert_allclose(mean, want, rtol=3e-4)
    assert math.isfinite(float(mean[0]))
