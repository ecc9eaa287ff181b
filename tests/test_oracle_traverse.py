"""The oracle's f32 traversal pinned by known answers and closed forms.

Watertight KATs restate packages/shared-bvh/src/__tests__/
opticalWatertightTriangle.test.ts:36-213.  The traversal is pinned by brute
force: the BVH closest hit equals the minimum over every triangle tested
with the same predicate, ties broken by the lowest packed triangle index
(the tie contract of the reference's left-first DFS with strict `<`).
"""
import numpy as np
import pytest

from oracle import oracle as O
from vitrum_amd import workloads as W


def next_f32(x):
    return (np.array([x], np.float32).view(np.uint32) + 1).view(np.float32)[0]


def wt(o, d, a, b, c, m=0.0):
    return O.watertight(o, d, a, b, c, m)


def test_planar_diagonal_exact_tie():
    f = wt([0, 0, 0], [0, 0, 1], [-1, -1, 1], [1, -1, 1], [1, 1, 1])
    s = wt([0, 0, 0], [0, 0, 1], [-1, -1, 1], [1, 1, 1], [-1, 1, 1])
    assert f["hit"] and s["hit"] and f["t"] == s["t"] == 1.0
    assert f["zero_edge_mask"] != 0 and s["zero_edge_mask"] != 0
    assert f["side"] == s["side"] == -1


def test_convex_edge_and_vertex_ties():
    ex = wt([2, 2, 0], [-1, -1, 0], [1, -1, -1], [1, 1, -1], [1, 1, 1])
    ey = wt([2, 2, 0], [-1, -1, 0], [-1, 1, -1], [1, 1, 1], [1, 1, -1])
    assert ex["hit"] and ey["hit"] and ex["t"] == ey["t"] == 1.0 and ex["side"] == ey["side"] == 1
    faces = [([1, -1, -1], [1, 1, -1], [1, 1, 1]), ([-1, 1, -1], [1, 1, 1], [1, 1, -1]),
             ([-1, -1, 1], [1, -1, 1], [1, 1, 1])]
    hits = [wt([2, 2, 2], [-1, -1, -1], *f) for f in faces]
    assert all(h["hit"] and h["zero_edge_mask"] >= 3 for h in hits)
    assert len({float(h["t"]) for h in hits}) == 1 and {h["side"] for h in hits} == {1}


def test_silhouette_edge_tangency_sides():
    f = wt([2, 0, 0], [-1, 1, 0], [1, -1, -1], [1, 1, -1], [1, 1, 1])
    b = wt([2, 0, 0], [-1, 1, 0], [-1, 1, -1], [1, 1, 1], [1, 1, -1])
    assert (f["side"], b["side"]) == (1, -1) and f["t"] == b["t"] == 1.0


def test_adjacent_f32_layers_distinct():
    inner_t = next_f32(1.0)
    outer = wt([0, 0, 0], [0, 0, 1], [-1, -1, 1], [1, -1, 1], [0, 1, 1])
    inner = wt([0, 0, 0], [0, 0, 1], [-1, -1, inner_t], [1, -1, inner_t], [0, 1, inner_t], outer["t"])
    assert outer["hit"] and inner["hit"] and outer["t"] == 1.0 and inner["t"] == inner_t


def test_shared_edge_point_is_canonical():
    base = 67_108_864.0
    ea = [base - 64, -base - 32, 1]
    eb = [base + 128, -base + 64, 1]
    o = [base, -base, 0]
    na = wt(o, [0, 0, 1], ea, eb, [base - 8_192, -base + 16_777_216, 1])
    nb = wt(o, [0, 0, 1], eb, ea, [base + 33_554_432, -base - 4_096, 1])
    assert na["hit"] and nb["hit"] and na["zero_edge_mask"] and nb["zero_edge_mask"]
    assert np.array_equal(na["point"], nb["point"]) and na["t"] == nb["t"]
    nz = next_f32(1.0)
    lifted = wt(o, [0, 0, 1], [ea[0], ea[1], nz], [eb[0], eb[1], nz],
                [base - 8_192, -base + 16_777_216, nz], na["t"])
    assert lifted["hit"] and lifted["t"] == nz and lifted["point"][2] == nz


def test_exact_edge_zeros_and_exclusive_min_t():
    e = wt([0, 0, 0], [0, 0, 1], [0, 0, 1], [1, 0, 1], [0, 1, 1], 0.0)
    assert e["hit"] and e["zero_edge_mask"] == 6
    i = wt([0.5, np.float32(0.49999994), 0], [0, 0, 1], [0, 0, 1], [1, 0, 1], [0, 1, 1])
    assert i["hit"] and i["zero_edge_mask"] == 0
    assert not wt([0, 0, 0], [0, 0, 1], [-1, -1, 1], [1, -1, 1], [0, 1, 1], 1.0)["hit"]
    n = next_f32(1.0)
    after = wt([0, 0, 0], [0, 0, 1], [-1, -1, n], [1, -1, n], [0, 1, n], 1.0)
    assert after["hit"] and after["t"] == n


def test_overflow_safe_side():
    h = wt([0, 0, 0], [1, 0, 0], [1, 0, 0], [1e30, 1e10, 0], [1, 0, 1e10])
    assert h["hit"] and h["t"] == 1.0 and h["side"] == -1


def test_interior_hit_matches_closed_form():
    # Axis-aligned triangle in z = 2: exact t and barycentrics are representable.
    h = wt([0.25, 0.25, 0], [0, 0, 1], [0, 0, 2], [1, 0, 2], [0, 1, 2])
    assert h["hit"] and h["t"] == 2.0
    np.testing.assert_array_equal(h["bary"], np.array([0.5, 0.25, 0.25], np.float32))
    np.testing.assert_array_equal(h["point"], np.array([0.25, 0.25, 2], np.float32))


# ---- traversal vs brute force ------------------------------------------------

def single_mesh_scene(seed, n_tri=200):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-1, 1, size=(n_tri * 3, 3)).astype(np.float32)
    nrm = np.tile([0, 0, 1], (n_tri * 3, 1)).astype(np.float32)
    return [dict(kind="mesh", positions=pos.ravel(), normals=nrm.ravel(),
                 indices=np.arange(n_tri * 3, dtype=np.uint32), material_id=0)]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_blas_traversal_equals_brute_force(seed):
    """Identity instance: the BVH result equals the brute-force minimum of the
    watertight predicate over all triangles (lowest index on exact ties)."""
    pack = O.pack_scene(single_mesh_scene(seed))
    ps = O.PackedScene(pack)
    rng = np.random.default_rng(100 + seed)
    n = 300
    rays = np.zeros(n, O.RAY_DTYPE)
    rays["o"] = rng.uniform(-2, 2, size=(n, 3))
    d = rng.normal(size=(n, 3))
    rays["d"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays["tmin"] = 1e-4
    rays["tmax"] = W.TMAX_INFINITY
    hits = O.trace_closest(ps, rays, threads=2)
    pos = pack["positions"].reshape(-1, 4)[:, :3]
    idx = pack["indices"].reshape(-1, 4)[:, :3]
    for i in range(n):
        best_t, best_tri = np.inf, None
        for t in range(idx.shape[0]):
            h = wt(rays["o"][i], rays["d"][i], pos[idx[t, 0]], pos[idx[t, 1]], pos[idx[t, 2]],
                   rays["tmin"][i])
            if h["hit"] and h["t"] > rays["tmin"][i] and h["t"] < best_t:
                best_t, best_tri = h["t"], t
        if best_tri is None:
            assert hits["inst"][i] == 0xFFFFFFFF
        else:
            assert hits["prim"][i] == best_tri, i
            # The reported t is re-derived from the reconstructed hit point
            # (dot(L2W * point - o, d)), which carries the f32 barycentric
            # error of the predicate; it agrees with the predicate's t closely.
            assert abs(float(hits["t"][i]) - float(best_t)) <= 2e-5 * max(1.0, abs(float(best_t)))


def test_any_hit_agrees_with_closest_on_opaque_scene():
    pack = O.pack_scene(single_mesh_scene(5, 400))
    ps = O.PackedScene(pack)
    rng = np.random.default_rng(9)
    n = 2000
    rays = np.zeros(n, O.RAY_DTYPE)
    rays["o"] = rng.uniform(-2, 2, size=(n, 3))
    d = rng.normal(size=(n, 3))
    rays["d"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays["tmin"] = 1e-4
    rays["tmax"] = W.TMAX_INFINITY
    hits = O.trace_closest(ps, rays)
    occ = O.trace_any(ps, rays)
    assert np.array_equal(occ.astype(bool), hits["inst"] != 0xFFFFFFFF)
    # With tmax just beyond the closest hit, any-hit still reports occlusion;
    # with tmax at the hit distance (exclusive) the nearer set is empty.
    hit = hits["inst"] != 0xFFFFFFFF
    r2 = rays[hit].copy()
    r2["tmax"] = hits["t"][hit]
    occ2 = O.trace_any(ps, r2)
    assert occ2.sum() <= hit.sum()


def test_fp64_parameters_close_on_cornell():
    sc = W.cornell_scene()
    pack = O.pack_scene([dict(p, material_id=i) for i, p in enumerate(sc["primitives"])])
    ps = O.PackedScene(pack)
    rays, _ = W.primary_rays(sc["camera"], 128, 128, scene_radius=sc["scene_radius"])
    hits = O.trace_closest(ps, rays)
    f = O.fp64_hit_params(ps, rays, hits)
    m = hits["inst"] != 0xFFFFFFFF
    assert m.mean() > 0.95
    rel = np.abs(hits["t"][m] - f[m, 0]) / np.abs(f[m, 0])
    # f32 traversal vs binary64 closed form: 99% within 1e-6 relative (the
    # 99.9th percentile sits near 1.04e-6), every hit within 1e-5.
    assert np.quantile(rel, 0.99) <= 1e-6 and rel.max() <= 1e-5
    # The reference's f32 watertight barycentrics carry edge-function
    # cancellation on small, distant triangles: bounded here at 2e-3 absolute.
    assert np.abs(hits["u"][m] - f[m, 1]).max() <= 2e-3
    assert np.abs(hits["v"][m] - f[m, 2]).max() <= 2e-3


# ---- the path tracer's closest-hit and visibility wrappers (a8) ------------


def _quad(y, up=True, material_id=0, transform=None):
    """A unit quad in the plane y = const, wound so that its normal is +y
    (up) or -y."""
    pos = np.array([[0, y, 0], [1, y, 0], [1, y, 1], [0, y, 1]], np.float32)
    idx = np.array([0, 2, 1, 0, 3, 2] if up else [0, 1, 2, 0, 2, 3], np.uint32)
    p = dict(kind="mesh", positions=pos.ravel(), normals=np.tile([0, 1 if up else -1, 0], 4).astype(np.float32),
             indices=idx, material_id=material_id)
    if transform is not None:
        p["transform"] = np.asarray(transform, np.float32)
    return p


def _sided_scene(mats, quads):
    pack = O.pack_scene(quads)
    ps = O.PackedScene(pack)
    return ps, O.single_sided_flags(ps, mats)


def _vertical_rays(ys, up):
    r = np.zeros(len(ys), O.RAY_DTYPE)
    r["o"] = [[0.3, y, 0.6] for y in ys]
    r["d"] = [0, 1, 0] if up else [0, -1, 0]
    r["tmin"] = 1e-4
    r["tmax"] = W.TMAX_INFINITY
    return r


def test_single_sided_back_faces_are_skipped():
    """materialAcceptsSidedHit: a single-sided (default) quad facing up at
    y = 1 over a single-sided floor facing up at y = 0.  From above, the
    quad's front face is the hit; from between them going up, the quad's back
    face is not a surface (miss); from below going up both back faces are
    skipped.  A double-sided quad is hit from either side.  The raw
    traversal reports the geometric closest hit in every case."""
    mats = [{}, {}]
    ps, ss = _sided_scene(mats, [_quad(0.0, True, 0), _quad(1.0, True, 1)])
    assert ss.tolist() == [1, 1, 1, 1]
    down = _vertical_rays([2.0], up=False)
    h = O.trace_closest_sided(ps, ss, down)
    assert h["t"][0] == np.float32(1.0) and h["inst"][0] == 1
    up = _vertical_rays([0.5, -1.0], up=True)
    h = O.trace_closest_sided(ps, ss, up)
    assert h["inst"].tolist() == [0xFFFFFFFF, 0xFFFFFFFF]
    raw = O.trace_closest(ps, up)
    assert raw["t"].tolist() == [0.5, 1.0]
    # visibility: occluded only through front faces
    vis_up = O.trace_occluded_sided(ps, ss, up)
    assert vis_up.tolist() == [0, 0]
    assert O.trace_occluded_sided(ps, ss, down).tolist() == [1]
    # double-sided upper quad: hit from below too
    ps2, ss2 = _sided_scene([{}, {"doubleSided": True}], [_quad(0.0, True, 0), _quad(1.0, True, 1)])
    assert ss2.tolist() == [1, 1, 0, 0]
    h = O.trace_closest_sided(ps2, ss2, up)
    assert h["inst"].tolist() == [1, 1] and h["t"].tolist() == [0.5, 2.0]
    assert O.trace_occluded_sided(ps2, ss2, up).tolist() == [1, 1]
    # with no flags every hit is a surface
    np.testing.assert_array_equal(O.trace_closest_sided(ps, np.zeros_like(ss), up)["t"], raw["t"])


def test_front_face_follows_the_instance_orientation():
    """intersection.wgsl.ts:45: an orientation-reversing instance flips the
    local front face.  The quad facing up, mirrored in x, is seen from above
    through its back face: the single-sided hit is skipped."""
    mirror = np.eye(4, dtype=np.float32)
    mirror[0, 0] = -1.0
    assert O.orientation_sign(mirror.T.ravel()) == -1
    assert O.orientation_sign(np.eye(4, dtype=np.float32).ravel()) == 1
    ps, ss = _sided_scene([{}], [_quad(1.0, True, 0, transform=mirror.T.ravel())])
    down = _vertical_rays([2.0], up=False)
    down["o"] = [[-0.3, 2.0, 0.6]]
    assert O.trace_closest(ps, down)["t"][0] == np.float32(1.0)
    assert O.trace_closest_sided(ps, ss, down)["inst"][0] == 0xFFFFFFFF
    up = _vertical_rays([0.0], up=True)
    up["o"] = [[-0.3, 0.0, 0.6]]
    assert O.trace_closest_sided(ps, ss, up)["t"][0] == np.float32(1.0)


def test_sided_cursor_is_exclusive():
    """The re-trace starts at the skipped hit's distance, exclusive: a front
    face coplanar with a skipped back face (same t) is not reported -- the
    reference's loop steps past every surface at that distance."""
    ps, ss = _sided_scene([{}, {}], [_quad(1.0, False, 0), _quad(1.0, True, 1)])
    up = _vertical_rays([0.0], up=True)
    raw = O.trace_closest(ps, up)
    assert raw["t"][0] == np.float32(1.0)
    h = O.trace_closest_sided(ps, ss, up)
    # the lower-indexed first hit is the back face of quad 1 (normal +y seen
    # from below) or the front face of quad 0 (normal -y): whichever the raw
    # traversal reports first decides
    if raw["inst"][0] == 0:
        assert h["inst"][0] == 0
    else:
        assert h["inst"][0] == 0xFFFFFFFF
