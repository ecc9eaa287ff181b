"""GPU wavefront path tracer against the binary64 CPU integrator and closed forms.

Parity contract (same sampler seeds on both sides):
  * path match: >= 99.9% of paths take identical discrete decisions (hit
    triangles, light picks, lobes, roulette), compared by per-path hashes;
  * matched paths: per-pixel radiance within 1e-4 relative + 1e-7 absolute
    for >= 99.9% of pixels (f32 device shading against binary64; both sides
    generate rays -- shading normal, sampled direction, next origin -- with
    the reference's f32 arithmetic, so every vertex, and the light tree's
    selection pdf there, is the same);
  * every pixel of a multi-frame mean: |mu_gpu - mu_cpu| <= 2e-4 |mu_cpu| +
    4 sqrt(var / N) with >= 99.9% of pixels passing.
"""

import numpy as np
import pytest

from vitrum_amd import workloads as W

from oracle import oracle as O
from render_cases import SEED, cpu_frames, directional_albedo_at_normal, materials_of, plane_under_sky

pytestmark = pytest.mark.gpu


def gpu_frames(engine, scene, res, frames, max_bounces=8):
    engine.set_scene(scene)
    engine.reset()
    ivp = W.inv_view_proj(scene["camera"], res, res)
    for f in range(frames):
        engine.render_frame(ivp, res, res, f, SEED, max_bounces)
    return engine.capture(res, res)[..., :3].astype(np.float64), engine.capture_signatures(res, res)


def check_single_frame(engine, scene, res, sampling=0, light_tree=True):
    gpu, gsig = gpu_frames(engine, scene, res, 1)
    info = engine.render_info()
    bias, _ = W.ray_bias(scene["scene_radius"])
    assert info["ray_bias"] == pytest.approx(float(bias), rel=1e-6)
    cpu, csig = cpu_frames(scene, res, 1, info["ray_bias"], sampling=sampling, light_tree=light_tree)
    cpu, csig = cpu[0], csig[0]
    match = gsig == csig
    rate = match.mean()
    assert rate >= 0.999, f"path match rate {rate:.5f}"
    diff = np.abs(gpu - cpu)
    tol = 1e-4 * np.abs(cpu) + 1e-7
    ok = (diff <= tol).all(axis=2) | ~match
    rel = (diff / (np.abs(cpu) + 1e-6)).max(axis=2)[match]
    print(f"path match {rate:.5f}; matched pixels within 1e-4: {np.mean(rel <= 1e-4):.5f}, "
          f"2e-4: {np.mean(rel <= 2e-4):.5f}; max relative {rel.max():.3g}", flush=True)
    assert ok.mean() >= 0.999, f"matched-path radiance agreement {ok.mean():.5f}"
    return gpu, cpu, rate


def test_cornell_paths_match_fp64_integrator(engine):
    sc = W.cornell_scene()
    gpu, cpu, rate = check_single_frame(engine, sc, 64)
    assert gpu.mean() > 0.05  # lit, not black


def test_spheres_paths_match_fp64_integrator(engine):
    sc = W.spheres_scene(instances=48, lights=4)
    gpu, cpu, rate = check_single_frame(engine, sc, 64)
    assert gpu.mean() > 0.01


@pytest.fixture(scope="module")
def sobol_engine():
    from vitrum_amd import native
    eng = native.Engine(0, sampling=1)
    yield eng
    eng.close()


def test_sobol_sampler_paths_match_fp64_integrator(sobol_engine):
    """Owen-scrambled Sobol mode: the same path-matching contract; the
    sampler's words are those of the oracle's pinned Sobol streams."""
    sc = W.cornell_scene()
    gpu, cpu, rate = check_single_frame(sobol_engine, sc, 64, sampling=1)
    assert gpu.mean() > 0.05
    sc = W.spheres_scene(instances=48, lights=4)
    check_single_frame(sobol_engine, sc, 48, sampling=1)


def test_light_tree_bytes_equal_oracle(engine):
    """The product's packed light tree (built on the host at setScene) equals
    the oracle's buildLightTree + packLightTreeForGPU over the same lights,
    bit for bit: Cornell (quad light, no sky: two triangle leaves) and the
    sphere scene (light-sphere triangles plus the sky leaf)."""
    from oracle import oracle as O
    from render_cases import materials_of
    for sc in (W.cornell_scene(), W.spheres_scene(instances=16, lights=4)):
        engine.set_scene(sc)
        got = engine.light_tree()
        pack = O.pack_scene([dict(p, material_id=i) for i, p in enumerate(sc["primitives"])])
        em = O.build_emitters(pack, O.material_table(materials_of(sc)))
        sky = sc.get("sky", [0, 0, 0])
        want = O.scene_light_tree(em, sky if any(sky) else None)
        assert want is not None and got.shape == want.shape
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.fixture(scope="module")
def uniform_engine():
    from vitrum_amd import native
    eng = native.Engine(0, light_selection=1)
    yield eng
    eng.close()


def test_uniform_light_selection_paths_match(uniform_engine):
    """lightTreeImportanceSampling: false -- uniform light selection: the
    same path-matching contract against the integrator in uniform mode."""
    sc = W.spheres_scene(instances=48, lights=4)
    check_single_frame(uniform_engine, sc, 48, light_tree=False)
    assert uniform_engine.light_tree().shape[0] == 0


def test_cornell_multiframe_mean_within_statistical_bound(engine):
    sc = W.cornell_scene()
    res, frames = 48, 8
    gpu, _ = gpu_frames(engine, sc, res, frames)
    bias = engine.render_info()["ray_bias"]
    cpu_all, _ = cpu_frames(sc, res, frames, bias)
    mu = cpu_all.mean(axis=0)
    var = cpu_all.var(axis=0, ddof=1) + 1e-12
    bound = 2e-4 * np.abs(mu) + 4.0 * np.sqrt(2.0 * var / frames) + 1e-6
    ok = (np.abs(gpu - mu) <= bound).all(axis=2)
    assert ok.mean() >= 0.999, ok.mean()


@pytest.mark.parametrize("roughness,metallic", [(1.0, 0.0), (0.3, 0.0), (0.5, 1.0)])
def test_plane_under_constant_sky_converges_to_quadrature(engine, roughness, metallic):
    """One bounce to a constant sky: the NEE + BSDF-sampling + MIS estimator
    must converge to L * (directional albedo), the albedo taken from a dense
    quadrature of the BSDF (independent of every sampling routine)."""
    a, L = 0.5, np.array([1.0, 0.8, 0.6])
    sc = plane_under_sky(a, L.tolist(), roughness, metallic)
    res, frames = 32, 96
    img, _ = gpu_frames(engine, sc, res, frames)
    expect = L * directional_albedo_at_normal(a, roughness, metallic)
    mean = img.reshape(-1, 3).mean(axis=0)
    np.testing.assert_allclose(mean, expect, rtol=1e-2)


def test_directly_viewed_emitter_is_exact(engine):
    """A camera ray that hits an emitter's front face adds Le unweighted."""
    pos = np.array([-5, -5, 0, 5, -5, 0, 5, 5, 0, -5, 5, 0], np.float32)
    nrm = np.tile([0, 0, 1], 4).astype(np.float32)
    idx = np.array([0, 1, 2, 0, 2, 3], np.uint32)
    prim = {"kind": "mesh", "positions": pos, "normals": nrm, "indices": idx,
            "material": {"baseColor": [0, 0, 0], "roughness": 1, "metallic": 0,
                         "emissive": [1.0, 0.5, 0.25], "emissiveIntensity": 2.0}}
    cam = W.look_at_camera(eye=(0, 0, 3), target=(0, 0, 0), up=(0, 1, 0), vfov_deg=30)
    sc = {"primitives": [prim], "camera": cam, "scene_radius": 8.0, "sky": [0, 0, 0]}
    img, _ = gpu_frames(engine, sc, 16, 2, max_bounces=1)
    np.testing.assert_allclose(img.reshape(-1, 3), np.tile([2.0, 1.0, 0.5], (256, 1)), rtol=1e-6)


# ---- mesh-area lights: closed forms (same scenes as the oracle's pins) ------

def _image_bound(img, expect, rel=2e-3):
    v = img.reshape(-1, 3)
    return v.mean(axis=0), 4 * v.std(axis=0, ddof=1) / np.sqrt(v.shape[0]) + rel * np.abs(expect)


@pytest.mark.parametrize("roughness,metallic,tree,viewed_wall_emits,n_sub", [
    (1.0, 0.0, False, True, 2), (0.2, 1.0, False, True, 1), (0.3, 0.0, True, True, 2),
    (1.0, 0.0, True, False, 2)])
def test_emissive_enclosure_one_bounce(engine, uniform_engine, roughness, metallic, tree, viewed_wall_emits, n_sub):
    """Inside a closed box of inward single-sided emitters the one-bounce
    pixel is Le (viewed wall's emission) + Le rho(wo) exactly: light pick,
    sqrt(r1) triangle sampling, area pdf with the light cosine and both MIS
    sides on the device."""
    from render_cases import emissive_box
    le, a = np.array([1.0, 0.7, 0.4]), 0.6
    sc = emissive_box(le.tolist(), a, roughness, metallic, n_sub=n_sub, viewed_wall_emits=viewed_wall_emits)
    img, _ = gpu_frames(engine if tree else uniform_engine, sc, 64, 64, max_bounces=1)
    expect = le * (float(viewed_wall_emits) + directional_albedo_at_normal(a, roughness, metallic))
    mean, bound = _image_bound(img, expect)
    assert (np.abs(mean - expect) <= bound).all(), (mean, expect, bound)


@pytest.mark.parametrize("roughness,tree", [(1.0, True), (1.0, False), (0.4, True)])
def test_disc_light_over_plane_matches_quadrature(engine, uniform_engine, roughness, tree):
    """A 48-gon emitting disc over a floor seen 30 degrees off normal: the
    device's one-bounce estimate converges to the area quadrature (pinned to
    Lambert's polygon formula)."""
    from render_cases import disc_expected_radiance, disc_over_plane
    le, a = np.array([30.0, 20.0, 10.0]), 0.5
    sc = disc_over_plane(le.tolist(), a, roughness, 0.0)
    img, _ = gpu_frames(engine if tree else uniform_engine, sc, 64, 64, max_bounces=1)
    expect = disc_expected_radiance(sc, le, a, roughness, 0.0)
    mean, bound = _image_bound(img, expect)
    assert (np.abs(mean - expect) <= bound).all(), (mean, expect, bound)


# ---- present pass and auxiliary outputs --------------------------------------

def test_present_pass_matches_restatement(engine):
    """captureFrame({colorSpace: 'output'}) for every operator, both colour
    spaces and several exposures equals the present restatement (binary64
    operators, binary16 store) to one binary16 step (f32 device arithmetic
    can land on the other side of a rounding boundary)."""
    from oracle import present as P
    sc = W.cornell_scene()
    img, _ = gpu_frames(engine, sc, 32, 4)
    for mode, idx in P.TONEMAP_MODE_INDEX.items():
        for cs, csi in (("srgb", 0), ("linear", 1)):
            for e in (1.0, 0.37, 3.0):
                got = engine.capture_output(32, 32, idx, e, csi)
                want = P.present(img, mode, e, cs)
                assert (got[..., 3] == 1).all()
                step = np.maximum(np.spacing(want.astype(np.float16)).astype(np.float64), 2.0 ** -24)
                assert (np.abs(got[..., :3] - want) <= step).all(), (mode, cs, e)


def test_first_hit_aux_and_variance_match_oracle(engine):
    """normalDepth / albedo of frame 0 equal the integrator's first-hit f32
    values; the luminance variance over 8 frames equals the variance of the
    binary64 path samples within f32 accumulation error."""
    sc = W.spheres_scene(instances=48, lights=4)
    res = 48
    img, gsig = gpu_frames(engine, sc, res, 1)
    nd = engine.capture_aux(0, res, res)
    alb = engine.capture_aux(1, res, res)
    bias = engine.render_info()["ray_bias"]
    pack = O.pack_scene([dict(p, material_id=i) for i, p in enumerate(sc["primitives"])])
    ivp = W.inv_view_proj(sc["camera"], res, res)
    _, csig, aux = O.render(pack, materials_of(sc), ivp, res, res, frame_seed=SEED, frame_index=0, max_bounces=8,
                            ray_bias=bias, env_radiance=sc["sky"], aux=True)
    same = gsig == csig
    assert same.mean() >= 0.999
    assert np.array_equal(nd[same], aux[same][:, 0:4].astype(np.float32))
    assert np.array_equal(alb[same], aux[same][:, 4:8].astype(np.float32))
    assert (aux[..., 3] > 0).mean() > 0.1  # the spheres cover part of the view
    frames = 8
    gpu_frames(engine, sc, res, frames)
    var = engine.capture_aux(2, res, res)[..., 0].astype(np.float64)
    cpu, _ = cpu_frames(sc, res, frames, bias, pack=pack)
    lum = 0.2126 * cpu[..., 0] + 0.7152 * cpu[..., 1] + 0.0722 * cpu[..., 2]
    want = np.maximum(0.0, (lum ** 2).mean(axis=0) - lum.mean(axis=0) ** 2)
    scale = (lum ** 2).mean(axis=0)
    ok = np.abs(var - want) <= 1e-5 * scale + 1e-9
    assert ok.mean() >= 0.999, ok.mean()


def test_material_sorted_shading_gives_identical_images(engine):
    """vt_set_shade_sort: shading each workgroup's queue entries in
    material-sorted order changes which lanes shade which path, not any
    path's result: images and path signatures are bit-identical."""
    for sc in (W.cornell_scene(), W.spheres_scene(instances=48, lights=4)):
        a, sa = gpu_frames(engine, sc, 48, 2)
        engine.set_shade_sort(True)
        try:
            b, sb = gpu_frames(engine, sc, 48, 2)
        finally:
            engine.set_shade_sort(False)
        assert np.array_equal(a, b) and np.array_equal(sa, sb)
