"""GPU parity of the cwbvh-closest traversal mode (bvhTraversal:
'cwbvh-closest', pt-webgpu pathTrace/intersection.wgsl.ts:576-838) through the
C ABI against the oracle's binary32 restatement of the same mode.

Bit-exact bar: prim, inst and the bits of t, u, v of every ray equal the
oracle's cwbvh-mode records.  The product builds its own wide forest from the
scene; equal records on every ray are the evidence that it walks the same
forest as the oracle's collapse.
"""
import numpy as np
import pytest

from oracle import oracle as O
from vitrum_amd import native as N
from vitrum_amd import workloads as W

pytestmark = pytest.mark.gpu


def assert_hits_equal(a, b, what=""):
    for k in ("prim", "inst", "t", "u", "v"):
        x = a[k].view(np.uint32)
        y = b[k].view(np.uint32)
        bad = np.nonzero(x != y)[0]
        assert bad.size == 0, f"{what} {k}: {bad.size} mismatches, first {bad[:5]} {a[bad[:3]]} {b[bad[:3]]}"


def oracle_scene(scene):
    prims = [dict(p, material_id=i) for i, p in enumerate(scene["primitives"])]
    pack = O.pack_scene(prims)
    return pack, O.PackedScene(pack)


@pytest.fixture(scope="module")
def cw_engine():
    eng = N.Engine(0, bvh_traversal=1)
    yield eng
    eng.close()


@pytest.fixture(scope="module")
def cornell_cw():
    sc = W.cornell_scene()
    pack, ps = oracle_scene(sc)
    rays, st = W.primary_rays(sc["camera"], 512, 512, scene_radius=sc["scene_radius"])
    hits = O.trace_closest(ps, rays)
    bounce = W.bounce_rays(pack, rays, hits, st, sc["scene_radius"])
    cw, roots = O.scene_cwbvh(ps)
    return sc, ps, np.concatenate([rays, bounce]), cw, roots


def test_cornell_cwbvh_bit_exact(cw_engine, cornell_cw):
    sc, ps, rays, cw, roots = cornell_cw
    print(f"cwbvh cornell: {rays.size} rays", flush=True)
    cw_engine.set_scene(sc)
    got = cw_engine.trace_closest(rays)
    assert_hits_equal(got, O.trace_closest_cwbvh(ps, cw, roots, rays), "cornell cwbvh")
    assert np.mean(got["inst"] != 0xFFFFFFFF) > 0.8  # bounce rays leave through the open front


def test_cwbvh_edge_case_rays(cw_engine, cornell_cw):
    """Axis-parallel directions, tMin / tMax windows (finite, the reference's
    INFINITY constant and +inf) and grid-line rays on shared edges."""
    sc, ps, _, cw, roots = cornell_cw
    cw_engine.set_scene(sc)
    rng = np.random.default_rng(7)
    n = 50000
    r = np.zeros(n, W.RAY_DTYPE)
    r["o"] = rng.uniform(-0.99, 0.99, size=(n, 3)) + np.array([0, 1, 0])
    d = rng.normal(size=(n, 3))
    d[: n // 3, rng.integers(0, 3, n // 3)] = 0.0
    d[n // 3: n // 2] = np.eye(3)[rng.integers(0, 3, n // 2 - n // 3)] * rng.choice([-1, 1], (n // 2 - n // 3, 1))
    d /= np.maximum(np.linalg.norm(d, axis=1, keepdims=True), 1e-30)
    r["d"] = d.astype(np.float32)
    r["tmin"] = rng.choice([0.0, 1e-4, 0.5], n).astype(np.float32)
    r["tmax"] = rng.choice([W.TMAX_INFINITY, np.float32(0.3), np.float32(np.inf)], n)
    assert_hits_equal(cw_engine.trace_closest(r), O.trace_closest_cwbvh(ps, cw, roots, r), "edge cases")
    g = np.zeros(4096, W.RAY_DTYPE)
    gx = (np.arange(64) / 32.0 - 1.0).astype(np.float32)
    xx, zz = np.meshgrid(gx, gx)
    g["o"] = np.stack([xx.ravel(), np.full(4096, 1.0), zz.ravel()], 1)
    g["d"] = np.array([0, -1, 0], np.float32)
    g["tmax"] = W.TMAX_INFINITY
    assert_hits_equal(cw_engine.trace_closest(g), O.trace_closest_cwbvh(ps, cw, roots, g), "grid-line rays")
    # any-hit rays stay on the binary traversal
    np.testing.assert_array_equal(cw_engine.trace_any(r), O.trace_any(ps, r))


def test_spheres_cwbvh_bit_exact(cw_engine):
    """The 1M-triangle, 264-instance scene: a 262,144-ray sample of the
    1024x1024 primary + bounce rays."""
    sc = W.spheres_scene()
    pack, ps = oracle_scene(sc)
    rays, st = W.primary_rays(sc["camera"], 1024, 1024, scene_radius=sc["scene_radius"])
    ph = O.trace_closest(ps, rays)
    bounce = W.bounce_rays(pack, rays, ph, st, sc["scene_radius"])
    allr = np.concatenate([rays, bounce])
    pick = np.random.default_rng(5).choice(allr.size, 262144, replace=False)
    sample = allr[np.sort(pick)]
    cw, roots = O.scene_cwbvh(ps)
    print(f"cwbvh spheres: {sample.size} rays", flush=True)
    cw_engine.set_scene(sc)
    got = cw_engine.trace_closest(sample)
    assert_hits_equal(got, O.trace_closest_cwbvh(ps, cw, roots, sample), "spheres cwbvh")


def test_switching_modes(cornell_cw):
    """vt_set_bvh_traversal switches an engine between the modes; mode 2 is
    rejected."""
    sc, ps, rays, cw, roots = cornell_cw
    eng = N.Engine(0)
    try:
        eng.set_scene(sc)
        sub = rays[:65536]
        eng.set_bvh_traversal(1)
        assert_hits_equal(eng.trace_closest(sub), O.trace_closest_cwbvh(ps, cw, roots, sub), "mode 1")
        eng.set_bvh_traversal(0)
        assert_hits_equal(eng.trace_closest(sub), O.trace_closest(ps, sub), "mode 0")
        with pytest.raises(N.VtError):
            eng.set_bvh_traversal(2)
    finally:
        eng.close()


def test_cwbvh_render_matches_binary_render(cw_engine):
    """The path tracer's extension rays in the cwbvh-closest mode: on the
    Cornell box every proposal re-verifies to the binary record, so the frame
    equals the binary mode's frame."""
    sc = W.cornell_scene()
    eng = N.Engine(0)
    try:
        res = 128
        ivp = W.inv_view_proj(sc["camera"], res, res)
        out = []
        for e in (eng, cw_engine):
            e.set_scene(sc)
            e.reset()
            for f in range(2):
                e.render_frame(ivp, res, res, f, 3)
            out.append((e.capture(res, res), e.capture_signatures(res, res)))
        (a, sa), (b, sb) = out
        same = np.all(a.view(np.uint32) == b.view(np.uint32), axis=-1)
        print(f"cwbvh render: {same.mean():.6f} of pixels bit-identical", flush=True)
        assert same.mean() >= 0.999
        assert np.mean(sa == sb) >= 0.999
    finally:
        eng.close()
