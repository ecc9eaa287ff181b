"""GPU traversal parity: the HIP kernels through the C ABI against the oracle.

Bit-exact bar: prim, inst and the bits of t, u, v of every ray equal the
oracle's (which follows the reference's f32 traversal semantics).
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
def cornell(engine):
    sc = W.cornell_scene()
    engine.set_scene(sc)
    pack, ps = oracle_scene(sc)
    rays, st = W.primary_rays(sc["camera"], 512, 512, scene_radius=sc["scene_radius"])
    hits = O.trace_closest(ps, rays)
    bounce = W.bounce_rays(pack, rays, hits, st, sc["scene_radius"])
    return sc, pack, ps, rays, hits, bounce


def test_fp32_divide_and_sqrt_round_like_host(engine):
    rng = np.random.default_rng(0)
    a = np.concatenate([rng.uniform(-1e3, 1e3, 200000), rng.uniform(0, 1e-38, 1000),
                        np.float32(2.0) ** rng.integers(-140, 120, 5000)]).astype(np.float32)
    b = np.concatenate([rng.uniform(-1e3, 1e3, 200000), rng.uniform(1e-3, 1e3, 1000),
                        rng.uniform(0.5, 3.0, 5000)]).astype(np.float32)
    q, r = engine.selftest_fp32(a, b)
    with np.errstate(all="ignore"):
        np.testing.assert_array_equal(q.view(np.uint32), (a / b).view(np.uint32))
        ra = np.abs(a)
        q2, r2 = engine.selftest_fp32(ra, b)
        np.testing.assert_array_equal(r2.view(np.uint32), np.sqrt(ra).view(np.uint32))


def test_cornell_primary_closest_bit_exact(engine, cornell):
    sc, pack, ps, rays, hits, _ = cornell
    engine.set_scene(sc)
    got = engine.trace_closest(rays)
    assert_hits_equal(got, hits, "cornell primary")
    assert np.mean(hits["inst"] != 0xFFFFFFFF) > 0.95


def test_cornell_bounce_closest_bit_exact(engine, cornell):
    sc, pack, ps, _, _, bounce = cornell
    engine.set_scene(sc)
    assert_hits_equal(engine.trace_closest(bounce), O.trace_closest(ps, bounce), "cornell bounce")


def test_cornell_work_counts_equal_oracle(engine, cornell):
    sc, pack, ps, rays, _, bounce = cornell
    engine.set_scene(sc)
    for r in (rays[:65536], bounce[:65536]):
        h, c = engine.trace_closest_counts(r)
        ho, co = O.trace_closest(ps, r, counts=True)
        assert_hits_equal(h, ho, "counted")
        np.testing.assert_array_equal(c, co)


def test_cornell_any_hit_equals_oracle(engine, cornell):
    sc, pack, ps, rays, hits, bounce = cornell
    engine.set_scene(sc)
    np.testing.assert_array_equal(engine.trace_any(bounce), O.trace_any(ps, bounce))
    # Shadow-style segments: tmax at a fraction of the closest distance.
    hit = hits["inst"] != 0xFFFFFFFF
    seg = rays[hit].copy()
    seg["tmax"] = hits["t"][hit] * np.float32(0.999)
    np.testing.assert_array_equal(engine.trace_any(seg), O.trace_any(ps, seg))
    seg["tmax"] = hits["t"][hit] * np.float32(1.001)
    occ = engine.trace_any(seg)
    np.testing.assert_array_equal(occ, O.trace_any(ps, seg))
    assert occ.all()


def test_edge_case_rays(engine, cornell):
    sc, pack, ps, *_ = cornell
    engine.set_scene(sc)
    rng = np.random.default_rng(7)
    n = 50000
    r = np.zeros(n, W.RAY_DTYPE)
    r["o"] = rng.uniform(-0.99, 0.99, size=(n, 3)) + np.array([0, 1, 0])
    d = rng.normal(size=(n, 3))
    # a third of the rays have axis-parallel directions (exact zero components)
    d[: n // 3, rng.integers(0, 3, n // 3)] = 0.0
    d[n // 3: n // 2] = np.eye(3)[rng.integers(0, 3, n // 2 - n // 3)] * rng.choice([-1, 1], (n // 2 - n // 3, 1))
    d /= np.maximum(np.linalg.norm(d, axis=1, keepdims=True), 1e-30)
    r["d"] = d.astype(np.float32)
    r["tmin"] = rng.choice([0.0, 1e-4, 0.5], n).astype(np.float32)
    r["tmax"] = rng.choice([W.TMAX_INFINITY, np.float32(0.3), np.float32(np.inf)], n)
    assert_hits_equal(engine.trace_closest(r), O.trace_closest(ps, r), "edge cases")
    np.testing.assert_array_equal(engine.trace_any(r), O.trace_any(ps, r))
    # Rays on grid lines of the tessellated walls hit shared edges and vertices.
    g = np.zeros(4096, W.RAY_DTYPE)
    gx = (np.arange(64) / 32.0 - 1.0).astype(np.float32)
    xx, zz = np.meshgrid(gx, gx)
    g["o"] = np.stack([xx.ravel(), np.full(4096, 1.0), zz.ravel()], 1)
    g["d"] = np.array([0, -1, 0], np.float32)
    g["tmax"] = W.TMAX_INFINITY
    assert_hits_equal(engine.trace_closest(g), O.trace_closest(ps, g), "grid-line rays")


def test_empty_batch(engine, cornell):
    engine.set_scene(cornell[0])
    assert engine.trace_closest(np.zeros(0, W.RAY_DTYPE)).size == 0
    assert engine.trace_any(np.zeros(0, W.RAY_DTYPE)).size == 0


def test_device_pointer_entry_points(engine, cornell):
    sc, pack, ps, rays, hits, _ = cornell
    engine.set_scene(sc)
    n = rays.size
    d_rays = engine.alloc(rays.nbytes)
    d_hits = engine.alloc(n * W.HIT_DTYPE.itemsize)
    try:
        engine.h2d(d_rays, rays)
        engine.trace_closest_device(d_rays, n, d_hits)
        engine.synchronize()
        out = np.zeros(n, W.HIT_DTYPE)
        engine.d2h(out, d_hits)
        assert_hits_equal(out, hits, "device entry")
        assert engine.stats()["last_trace_ms"] > 0
    finally:
        engine.free(d_rays)
        engine.free(d_hits)


def test_golden_scene_routes(engine):
    from tests.test_oracle_build import GOLDEN_SCENES
    prims = GOLDEN_SCENES["multiMesh"]()
    scene = {"primitives": [dict(p, material={}) for p in prims], "scene_radius": 10.0}
    engine.set_scene(scene)
    pack = O.pack_scene(prims)
    ps = O.PackedScene(pack)
    spec = [([-1, .5, .5], [1, 0, 0], 2), ([2, .5, .5], [1, 0, 0], 2.5), ([5.2, .2, 1], [0, 0, -1], 2),
            ([7.2, .2, 1], [0, 0, -1], 2), ([9.5, .5, .5], [1, 0, 0], 2)]
    r = np.zeros(len(spec), W.RAY_DTYPE)
    for i, (o, d, tmax) in enumerate(spec):
        r[i] = (o, 0.0, d, tmax)
    got = engine.trace_closest(r)
    assert_hits_equal(got, O.trace_closest(ps, r), "golden routes")
    assert got["inst"].tolist() == [0, 1, 2, 3, 4]


def test_spheres_full_frame_bit_exact(engine):
    """The 1M-triangle scene at 1024x1024 primary + 1 bounce per hit."""
    sc = W.spheres_scene()
    engine.set_scene(sc)
    pack, ps = oracle_scene(sc)
    rays, st = W.primary_rays(sc["camera"], 1024, 1024, scene_radius=sc["scene_radius"])
    got = engine.trace_closest(rays)
    ref = O.trace_closest(ps, rays)
    assert_hits_equal(got, ref, "spheres primary")
    bounce = W.bounce_rays(pack, rays, ref, st, sc["scene_radius"])
    assert_hits_equal(engine.trace_closest(bounce), O.trace_closest(ps, bounce), "spheres bounce")
    np.testing.assert_array_equal(engine.trace_any(bounce), O.trace_any(ps, bounce))


def walk_agreement(got, ts):
    """Rays whose GPU hit record names the same (triangle, instance) -- or the
    same miss -- as the reference's binary64 wide-BVH CPU walk."""
    hw = got["inst"] != 0xFFFFFFFF
    ht = ts["hit"] == 1
    return (hw == ht) & (~hw | ((got["prim"] == ts["prim"]) & (got["inst"] == ts["inst"])))


def test_cornell_ids_match_reference_cpu_walk(engine, cornell):
    """Hit records against the reference's own CPU closest hit
    (intersectCompressedWideBvhFirstHit over the collapsed wide BVH,
    compressedWideBvh.ts:754, composed with tlasIntersect): identical
    (triangle, instance) ids for at least 99.9% of the 512x512 primary +
    bounce rays (edge-grazing rays may differ: the walk's 1e-6 barycentric
    tolerance admits both neighbours of a shared edge), and t within 1e-6
    relative of the walk's binary64 t for 99% of primary rays."""
    sc, pack, ps, rays, hits, bounce = cornell
    engine.set_scene(sc)
    allr = np.concatenate([rays, bounce])
    got = engine.trace_closest(allr)
    cw, roots = O.scene_cwbvh(ps)
    ts = O.cwbvh_trace_closest(ps, cw, roots, allr)
    same = walk_agreement(got, ts)
    print(f"cornell: {same.mean():.6f} of {allr.size} rays agree with the CPU walk")
    assert same.mean() >= 0.999, int((~same).sum())
    prim = np.zeros(allr.size, bool)
    prim[:rays.size] = True
    m = prim & same & (ts["hit"] == 1)
    rel = np.abs(got["t"][m].astype(np.float64) - ts["t"][m]) / np.abs(ts["t"][m])
    assert np.quantile(rel, 0.99) <= 1e-6


def test_spheres_ids_match_reference_cpu_walk(engine):
    """The 1M-triangle, 256-instance scene: a 262,144-ray sample of the
    1024x1024 primary + bounce rays against the CPU wide-BVH walk."""
    sc = W.spheres_scene()
    engine.set_scene(sc)
    pack, ps = oracle_scene(sc)
    rays, st = W.primary_rays(sc["camera"], 1024, 1024, scene_radius=sc["scene_radius"])
    ph = engine.trace_closest(rays)
    bounce = W.bounce_rays(pack, rays, ph, st, sc["scene_radius"])
    allr = np.concatenate([rays, bounce])
    pick = np.random.default_rng(11).choice(allr.size, 262144, replace=False)
    sample = allr[np.sort(pick)]
    got = engine.trace_closest(sample)
    cw, roots = O.scene_cwbvh(ps)
    assert np.all(cw.status == 0)
    ts = O.cwbvh_trace_closest(ps, cw, roots, sample)
    same = walk_agreement(got, ts)
    print(f"spheres: {same.mean():.6f} of {sample.size} rays agree with the CPU walk")
    assert same.mean() >= 0.999, int((~same).sum())


def test_node_widths_agree(engine):
    """The 4-wide trees collapsed from the binary BVH and TLAS (default) and
    the binary trees give the same hit records on the 1M-triangle scene (a
    262,144-ray sample of primary + bounce rays), closest and any hit; the
    oracle comparison of the default width is in the tests above."""
    sc = W.spheres_scene()
    engine.set_scene(sc)
    pack, ps = oracle_scene(sc)
    rays, st = W.primary_rays(sc["camera"], 1024, 1024, scene_radius=sc["scene_radius"])
    ph = engine.trace_closest(rays)
    bounce = W.bounce_rays(pack, rays, ph, st, sc["scene_radius"])
    allr = np.concatenate([rays, bounce])
    sample = allr[np.sort(np.random.default_rng(3).choice(allr.size, 262144, replace=False))]
    try:
        engine.set_node_width(2)
        h2, a2 = engine.trace_closest(sample), engine.trace_any(sample)
        for near in (0,):
            engine.set_trace_tuning(8, 16, near, 16)
            h2n = engine.trace_closest(sample)
            engine.set_node_width(4)
            assert_hits_equal(engine.trace_closest(sample), h2n, f"near_first={near}")
            engine.set_node_width(2)
        engine.set_trace_tuning(8, 16, 1, 12)
        engine.set_node_width(4)
        assert_hits_equal(engine.trace_closest(sample), h2, "width 4 vs 2")
        np.testing.assert_array_equal(engine.trace_any(sample), a2)
        with pytest.raises(N.VtError):
            engine.set_node_width(3)
    finally:
        engine.set_trace_tuning(8, 16, 1, 12)
        engine.set_node_width(4)
