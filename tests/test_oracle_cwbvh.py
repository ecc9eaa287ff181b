"""The reference's own CPU closest hit (8-wide compressed BVH, binary64
Moller-Trumbore) restated in the oracle, pinned by the known answers of
packages/shared-bvh/src/__tests__/compressedWideBvh.test.ts and
cwbvhTMinParity.test.ts, by brute force, and compared with the f32
watertight traversal that the GPU path reproduces.
"""
import numpy as np
import pytest

from oracle import oracle as O
from vitrum_amd import workloads as W

CHILD_NODE, CHILD_LEAF = 1, 2


def make_grid(size):
    """compressedWideBvh.test.ts:23-53: a (size x size) quad grid in z = 0."""
    pos = []
    for y in range(size + 1):
        for x in range(size + 1):
            pos += [x, y, 0, 0]
    idx, mats = [], []
    for y in range(size):
        for x in range(size):
            v0 = y * (size + 1) + x
            v1, v2, v3 = v0 + 1, v0 + size + 2, v0 + size + 1
            idx += [v0, v1, v2, 0, v0, v2, v3, 0]
            mats += [(x + y) & 3, (x + y + 1) & 3]
    return np.array(pos, np.float32), np.array(idx, np.uint32), np.array(mats, np.uint32)


def blas_scene(positions, indices, mats, max_leaf=4):
    """A single BLAS (binary build of the oracle) viewed as a packed scene
    with one identity instance, plus its wide tree."""
    b = O.build_blas(positions, indices, mats, max_leaf=max_leaf)
    tri = indices.size // 4
    cw = O.Cwbvh(b["nodes"], [0], tri_count=tri)
    return b, cw


def mt_brute(o, d, a, b, c):
    """The test's own brute-force predicate (compressedWideBvh.test.ts:75-96)."""
    o, d, a, b, c = (np.asarray(v, np.float64) for v in (o, d, a, b, c))
    e1, e2 = b - a, c - a
    n = np.cross(e1, e2)
    det = -np.dot(d, n)
    if abs(det) < 1e-5:
        return None
    ao = o - a
    dao = np.cross(ao, d)
    u = np.dot(e2, dao) / det
    v = -np.dot(e1, dao) / det
    t = np.dot(ao, n) / det
    if u < -1e-5 or v < -1e-5 or 1 - u - v < -1e-5 or t < 1e-5:
        return None
    return t


def single_instance_scene(positions, indices):
    """packSceneFromCore of one mesh (identity instance)."""
    p = positions.reshape(-1, 4)[:, :3]
    nrm = np.tile([0, 0, 1], (p.shape[0], 1)).astype(np.float32)
    pack = O.pack_scene([dict(kind="mesh", positions=p.ravel(), normals=nrm.ravel(),
                              indices=indices.reshape(-1, 4)[:, :3].ravel(), material_id=0)])
    return O.PackedScene(pack)


def rays_of(origins, dirs, tmin=1e-5, tmax=W.TMAX_INFINITY):
    r = np.zeros(len(origins), O.RAY_DTYPE)
    r["o"] = origins
    r["d"] = dirs
    r["tmin"] = tmin
    r["tmax"] = tmax
    return r


def test_empty_input_is_one_childless_root():
    """compressedWideBvh.test.ts:258-273."""
    b = O.build_blas(np.zeros(0, np.float32), np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    cw = O.Cwbvh(b["nodes"], [0], tri_count=0)
    assert cw.node_count == 1 and cw.child_count[0] == 0 and cw.status[0] == 1
    assert cw.node_bounds.shape == (1, 6) and cw.child_meta.shape == (1, 8, 3)


def test_single_triangle_is_one_leaf_child():
    """compressedWideBvh.test.ts:275-297."""
    pos = np.array([0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0], np.float32)
    idx = np.array([0, 1, 2, 0], np.uint32)
    _, cw = blas_scene(pos, idx, np.zeros(1, np.uint32))
    assert cw.node_count == 1 and cw.child_count[0] == 1 and cw.status[0] == 0
    assert cw.child_meta[0, 0].tolist() == [CHILD_LEAF, 0, 1]
    ps = single_instance_scene(pos, idx)
    cwp, roots = O.scene_cwbvh(ps)
    h = O.cwbvh_trace_closest(ps, cwp, roots, rays_of([[0.25, 0.25, 1]], [[0, 0, -1]]))
    assert h["hit"][0] == 1 and ps.pack["tri_source"][h["prim"][0]] == 0


def test_collapse_is_deterministic_and_bounded():
    """compressedWideBvh.test.ts:299-325 plus the collapse's invariants: at
    most 8 live children, every triangle in exactly one leaf, node bounds equal
    the collapsed binary node's bounds."""
    pos, idx, mats = make_grid(8)
    b, a = blas_scene(pos, idx, mats, max_leaf=1)
    _, c = blas_scene(pos, idx, mats, max_leaf=1)
    for k in ("node_bounds", "child_bounds", "child_meta", "child_count"):
        assert np.array_equal(getattr(a, k), getattr(c, k))
    tri = idx.size // 4
    seen = np.zeros(tri, np.int32)
    for n in range(a.node_count):
        cnt = a.child_count[n]
        assert 1 <= cnt <= 8
        assert np.all(a.child_meta[n, cnt:] == 0)
        for s in range(cnt):
            kind, off, num = a.child_meta[n, s]
            assert kind in (CHILD_NODE, CHILD_LEAF)
            if kind == CHILD_LEAF:
                seen[off:off + num] += 1
            else:
                assert off > n  # preorder allocation
    assert np.all(seen == 1)
    root_bits = b["nodes"].reshape(-1, 8)[0, :6].view(np.float32)
    assert np.array_equal(a.node_bounds[0], root_bits)
    # a tree: every wide node but the root is the child of exactly one slot;
    # with leaf size 1 the 128 triangles are 128 leaf slots and the 127
    # interior binary nodes are split among at most 127 wide nodes
    n_leaf = int(sum((a.child_meta[n, :a.child_count[n], 0] == CHILD_LEAF).sum() for n in range(a.node_count)))
    assert n_leaf == tri and int(a.child_count.sum()) == a.node_count - 1 + n_leaf
    assert a.node_count <= tri - 1


def test_quantized_child_bounds_are_conservative():
    """compressedWideBvh.test.ts:348-393."""
    pos, idx, mats = make_grid(4)
    b, cw = blas_scene(pos, idx, mats, max_leaf=1)
    P = pos.reshape(-1, 4)[:, :3].astype(np.float64)
    ri = b["reordered_indices"].reshape(-1, 4)
    for n in range(cw.node_count):
        for s in range(cw.child_count[n]):
            kind, off, num = cw.child_meta[n, s]
            mn, mx = cw.child_box(n, s)
            if kind == CHILD_NODE:
                exp_mn = cw.node_bounds[off, :3].astype(np.float64)
                exp_mx = cw.node_bounds[off, 3:].astype(np.float64)
            else:
                v = P[ri[off:off + num, :3].ravel()]
                exp_mn, exp_mx = v.min(axis=0), v.max(axis=0)
            assert np.all(np.array(mn) <= exp_mn + 1e-6) and np.all(np.array(mx) >= exp_mx - 1e-6)


def test_quantization_rounds_outward_by_one_step():
    """quantizeBound (compressedWideBvh.ts:197-204): floor-1 / ceil+1 of the
    unit offset times 65535, clamped; a child equal to its parent spans
    [0, 65535]."""
    pos, idx, mats = make_grid(4)
    _, cw = blas_scene(pos, idx, mats, max_leaf=1)
    for n in range(cw.node_count):
        pmn, pmx = cw.node_bounds[n, :3].astype(np.float64), cw.node_bounds[n, 3:].astype(np.float64)
        for s in range(cw.child_count[n]):
            kind, off, _ = cw.child_meta[n, s]
            if kind != CHILD_NODE:
                continue
            cmn, cmx = cw.node_bounds[off, :3].astype(np.float64), cw.node_bounds[off, 3:].astype(np.float64)
            for a in range(3):
                if pmx[a] == pmn[a]:
                    assert cw.child_bounds[n, s, a] == 0 and cw.child_bounds[n, s, 3 + a] == 65535
                    continue
                qlo = np.floor((cmn[a] - pmn[a]) / (pmx[a] - pmn[a]) * 65535) - 1
                qhi = np.ceil((cmx[a] - pmn[a]) / (pmx[a] - pmn[a]) * 65535) + 1
                assert cw.child_bounds[n, s, a] == int(np.clip(qlo, 0, 65535))
                assert cw.child_bounds[n, s, 3 + a] == int(np.clip(qhi, 0, 65535))


def test_stack_capacity_proof():
    """compressedWideBvh.test.ts:212-256: a chain 65 deep needs more than the
    64-entry traversal stack; a cycle fails the proof."""
    depth = 65
    count = np.zeros(depth * 2, np.uint32)
    meta = np.zeros((depth * 2, 8, 3), np.uint32)
    for level in range(depth):
        deep, sib = level * 2, level * 2 + 1
        count[sib] = 1
        meta[sib, 0, 0] = CHILD_LEAF
        if level + 1 < depth:
            count[deep] = 2
            meta[deep, 0] = [CHILD_NODE, sib, 0]
            meta[deep, 1] = [CHILD_NODE, deep + 2, 0]
        else:
            count[deep] = 1
            meta[deep, 0, 0] = CHILD_LEAF
    assert O.cwbvh_stack_entries(meta, count) > 64
    cyc = np.zeros((1, 8, 3), np.uint32)
    cyc[0, 0] = [CHILD_NODE, 0, 0]
    assert O.cwbvh_stack_entries(cyc, np.array([1], np.uint32)) == np.inf


def test_first_hit_matches_brute_force_on_grid():
    """compressedWideBvh.test.ts:395-415 (same rays, same brute force)."""
    pos, idx, mats = make_grid(5)
    ps = single_instance_scene(pos, idx)
    cw, roots = O.scene_cwbvh(ps)
    origins = [[0.25, 0.75, 3], [1.25, 2.75, 3], [4.25, 3.75, 3], [-0.5, 2, 3], [2, 5.5, 3]]
    h = O.cwbvh_trace_closest(ps, cw, roots, rays_of(origins, [[0, 0, -1]] * 5))
    P = pos.reshape(-1, 4)[:, :3]
    I = idx.reshape(-1, 4)[:, :3]
    for i, o in enumerate(origins):
        best, src = np.inf, -1
        for t in range(I.shape[0]):
            tt = mt_brute(o, [0, 0, -1], P[I[t, 0]], P[I[t, 1]], P[I[t, 2]])
            if tt is not None and tt < best:
                best, src = tt, t
        assert bool(h["hit"][i]) == (src >= 0)
        if src >= 0:
            assert abs(h["t"][i] - best) < 1e-6
            assert ps.pack["tri_source"][h["prim"][i]] == src


def test_open_interval_tmin_tmax():
    """cwbvhTMinParity.test.ts:31-80: a hit exactly at tMin or at tMax is
    rejected; just above tMin it is accepted."""
    z = 5.0
    pos = np.array([-1, -1, z, 0, 3, -1, z, 0, -1, 3, z, 0], np.float32)
    idx = np.array([0, 1, 2, 0], np.uint32)
    ps = single_instance_scene(pos, idx)
    cw, roots = O.scene_cwbvh(ps)
    o, d = [[0, 0, 0]], [[0, 0, 1]]
    assert O.cwbvh_trace_closest(ps, cw, roots, rays_of(o, d, tmin=z))["hit"][0] == 0
    h = O.cwbvh_trace_closest(ps, cw, roots, rays_of(o, d, tmin=z - 1e-4))
    assert h["hit"][0] == 1 and abs(h["t"][0] - z) < 1e-6
    assert O.cwbvh_trace_closest(ps, cw, roots, rays_of(o, d, tmin=0.0, tmax=z))["hit"][0] == 0


def test_random_soup_matches_exhaustive_moller_trumbore():
    """Closed form: the wide walk returns the minimum binary64 t over every
    triangle under the same predicate (no culling errors)."""
    rng = np.random.default_rng(4)
    n_tri = 300
    p = rng.uniform(-1, 1, size=(n_tri * 3, 3)).astype(np.float32)
    pos = np.concatenate([p, np.zeros((p.shape[0], 1), np.float32)], axis=1).ravel()
    idx = np.concatenate([np.arange(n_tri * 3, dtype=np.uint32).reshape(-1, 3),
                          np.zeros((n_tri, 1), np.uint32)], axis=1).ravel()
    ps = single_instance_scene(pos, idx)
    cw, roots = O.scene_cwbvh(ps)
    n = 400
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = rays_of(rng.uniform(-2, 2, size=(n, 3)), d.astype(np.float32), tmin=1e-4)
    h = O.cwbvh_trace_closest(ps, cw, roots, rays)
    # exhaustive: one-triangle-per-instance scene is not needed; walk all
    # triangles through single-triangle wide trees of the same scene
    P = ps.pack["positions"].reshape(-1, 4)[:, :3].astype(np.float64)
    I = ps.pack["indices"].reshape(-1, 4)[:, :3]
    for i in range(n):
        o = rays["o"][i].astype(np.float64)
        dd = rays["d"][i].astype(np.float64)
        best = np.inf
        for t in range(I.shape[0]):
            a, b, c = P[I[t, 0]], P[I[t, 1]], P[I[t, 2]]
            e1, e2 = b - a, c - a
            nn = np.cross(e1, e2)
            det = -np.dot(dd, nn)
            if abs(det) / (np.linalg.norm(nn) * np.linalg.norm(dd)) <= 1e-7:
                continue
            ao = o - a
            dao = np.cross(ao, dd)
            u, v = np.dot(e2, dao) / det, -np.dot(e1, dao) / det
            tt = np.dot(ao, nn) / det
            if u >= -1e-6 and v >= -1e-6 and 1 - u - v >= -1e-6 and tt > 1e-4 and tt < best:
                best = tt
        if np.isinf(best):
            assert h["hit"][i] == 0, i
        else:
            assert h["hit"][i] == 1 and abs(h["t"][i] - best) <= 1e-9 * max(1, best), i


@pytest.fixture(scope="module")
def cornell_rays():
    sc = W.cornell_scene()
    pack = O.pack_scene([dict(p, material_id=i) for i, p in enumerate(sc["primitives"])])
    ps = O.PackedScene(pack)
    prim, st = W.primary_rays(sc["camera"], 128, 128, scene_radius=sc["scene_radius"])
    ph = O.trace_closest(ps, prim)
    rays = np.concatenate([prim, W.bounce_rays(pack, prim, ph, st, sc["scene_radius"])])
    return ps, rays, prim.size


def test_watertight_traversal_agrees_with_reference_cpu_walk(cornell_rays):
    """The f32 watertight traversal (what the GPU kernels reproduce bit for
    bit) against the reference's binary64 wide-BVH walk on the 32,768-triangle
    box: the same (triangle, instance) for all but edge-grazing rays, where
    the walk's 1e-6 barycentric tolerance admits both neighbours; t within
    1e-6 relative of each other."""
    ps, rays, n_prim = cornell_rays
    wt = O.trace_closest(ps, rays)
    cw, roots = O.scene_cwbvh(ps)
    assert np.all(cw.status == 0)
    ts = O.cwbvh_trace_closest(ps, cw, roots, rays)
    hit_w = wt["inst"] != 0xFFFFFFFF
    hit_t = ts["hit"] == 1
    same = (hit_w == hit_t) & (~hit_w | ((wt["prim"] == ts["prim"]) & (wt["inst"] == ts["inst"])))
    # measured: all 32,768 rays agree
    assert same.mean() >= 0.999, (1 - same.mean(), int((~same).sum()))
    both = hit_w & hit_t
    # Where they differ the two candidates are at the same distance (an edge
    # or vertex shared by neighbours), not a culling difference.
    diff = ~same & hit_w & hit_t
    if diff.any():
        r2 = np.abs(wt["t"][diff].astype(np.float64) - ts["t"][diff]) / np.abs(ts["t"][diff])
        assert r2.max() <= 1e-5
    # t: the f32 traversal re-derives t from the reconstructed hit point; on
    # primary rays 99% are within 1e-6 relative of the binary64 walk (short
    # bounce segments carry more relative rounding, median still ~1e-7).
    rel = np.abs(wt["t"].astype(np.float64) - ts["t"]) / np.abs(ts["t"])
    prim_both = both.copy()
    prim_both[n_prim:] = False
    assert np.quantile(rel[prim_both], 0.99) <= 1e-6
    assert np.quantile(rel[both], 0.5) <= 1e-6
    # Two independent binary64 computations of the same hit -- the walk's
    # scale-normalised solve and the closed form on the watertight hit's
    # triangle -- agree to a few ulps of f32.
    f = O.fp64_hit_params(ps, rays, wt)
    rel64 = np.abs(f[same & both, 0] - ts["t"][same & both]) / np.abs(ts["t"][same & both])
    assert rel64.max() <= 1e-6


# ---- pt-webgpu's `cwbvh-closest` mode (binary32) ---------------------------


class _Forest:
    """A Cwbvh's arrays, copied so a test can corrupt them."""

    def __init__(self, cw):
        self.node_bounds = cw.node_bounds.copy()
        self.child_bounds = cw.child_bounds.copy()
        self.child_meta = cw.child_meta.copy()
        self.child_count = cw.child_count.copy()
        self.node_count = cw.node_count


def test_f32_mode_known_answer_single_triangle():
    """An axis-aligned ray through (0.25, 0.25) of a unit right triangle at
    z = 2: Moller-Trumbore and the watertight test both give t = 2 exactly,
    so the proposal is verified and the record is the binary mode's."""
    pos = np.array([0, 0, 2, 0, 1, 0, 2, 0, 0, 1, 2, 0], np.float32)
    idx = np.array([0, 1, 2, 0], np.uint32)
    ps = single_instance_scene(pos, idx)
    cw, roots = O.scene_cwbvh(ps)
    r = rays_of([[0.25, 0.25, 0.0], [0.25, 0.25, 3.0], [2.0, 2.0, 0.0]],
                [[0, 0, 1], [0, 0, 1], [0, 0, 1]])
    h = O.trace_closest_cwbvh(ps, cw, roots, r)
    assert h["prim"][0] == 0 and h["t"][0] == np.float32(2.0)
    assert h["u"][0] == np.float32(0.25) and h["v"][0] == np.float32(0.25)
    assert h["inst"][1] == 0xFFFFFFFF and h["inst"][2] == 0xFFFFFFFF  # behind / beside
    assert_records_equal(h, O.trace_closest(ps, r))


def test_f32_mode_rejects_grazing_ray():
    """A ray in the triangle's plane: the angular determinant test
    (|dot(rd, n)| <= 1e-7) rejects the proposal, and the watertight binary
    traversal misses as well."""
    pos = np.array([0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0], np.float32)
    idx = np.array([0, 1, 2, 0], np.uint32)
    ps = single_instance_scene(pos, idx)
    cw, roots = O.scene_cwbvh(ps)
    r = rays_of([[-1.0, 0.25, 0.0]], [[1, 0, 0]])
    assert O.trace_closest_cwbvh(ps, cw, roots, r)["inst"][0] == 0xFFFFFFFF


def assert_records_equal(a, b):
    for k in ("prim", "inst", "t", "u", "v"):
        np.testing.assert_array_equal(a[k].view(np.uint32), b[k].view(np.uint32), err_msg=k)


def test_f32_mode_agrees_with_binary_mode(cornell_rays):
    """Every proposal that survives the bit-exact watertight re-check is a
    watertight hit, so the mode returns the binary traversal's record except
    where a Moller-Trumbore proposal beats the exact closest triangle's
    (then its t can only be larger) -- measured: no such ray among the
    32,768 primary + bounce rays."""
    ps, rays, _ = cornell_rays
    cw, roots = O.scene_cwbvh(ps)
    got = O.trace_closest_cwbvh(ps, cw, roots, rays)
    ref = O.trace_closest(ps, rays)
    same = np.ones(rays.size, bool)
    for k in ("prim", "inst", "t", "u", "v"):
        same &= got[k].view(np.uint32) == ref[k].view(np.uint32)
    assert same.mean() >= 0.999, int((~same).sum())
    both = ~same & (got["inst"] != 0xFFFFFFFF) & (ref["inst"] != 0xFFFFFFFF)
    assert np.all(got["t"][both] >= ref["t"][both])


def test_f32_mode_falls_back_on_invalid_forest(cornell_rays):
    """intersection.wgsl.ts:576-700: an invalid wide layout (child count above
    8, a child of kind 0, a wide root outside the forest) answers the BLAS
    with the binary traversal, whose records are returned unchanged."""
    ps, rays, _ = cornell_rays
    rays = rays[:4096]
    cw, roots = O.scene_cwbvh(ps)
    ref = O.trace_closest(ps, rays)
    bad = _Forest(cw)
    bad.child_count[:] = 0xFFFFFFFF
    assert_records_equal(O.trace_closest_cwbvh(ps, bad, roots, rays), ref)
    bad = _Forest(cw)
    bad.child_meta[:, :, 0] = 0
    assert_records_equal(O.trace_closest_cwbvh(ps, bad, roots, rays), ref)
    far = np.full_like(roots, cw.node_count + 5)
    assert_records_equal(O.trace_closest_cwbvh(ps, cw, far, rays), ref)


def test_f32_mode_tmin_tmax_window():
    """cwbvhTMinParity.test.ts semantics in the f32 mode: hits at t <= tMin
    or t >= tMax are not reported, the next surface beyond tMin is."""
    pos, idx, mats = make_grid(4)
    # a second layer at z = 1 (translated copy of the grid)
    p2 = pos.reshape(-1, 4).copy()
    p2[:, 2] = 1.0
    pos2 = np.concatenate([pos.reshape(-1, 4), p2]).ravel()
    idx2 = np.concatenate([idx.reshape(-1, 4), idx.reshape(-1, 4) + np.array([25, 25, 25, 0], np.uint32)]).ravel()
    ps = single_instance_scene(pos2.astype(np.float32), idx2.astype(np.uint32))
    cw, roots = O.scene_cwbvh(ps)
    o = [[1.3, 2.7, -1.0]] * 4
    d = [[0, 0, 1]] * 4
    r = rays_of(o, d)
    r["tmin"] = np.array([0.0, 1.0, 1.5, 0.0], np.float32)
    r["tmax"] = np.array([W.TMAX_INFINITY, W.TMAX_INFINITY, W.TMAX_INFINITY, 1.0], np.float32)
    h = O.trace_closest_cwbvh(ps, cw, roots, r)
    # layers at t = 1 and t = 2: (0, inf) -> 1, (1, inf) -> 2, (1.5, inf) -> 2, (0, 1) -> miss
    assert h["t"][:3].tolist() == [1.0, 2.0, 2.0]
    assert h["inst"][3] == 0xFFFFFFFF
    assert_records_equal(h, O.trace_closest(ps, r))


def test_f32_mode_rechecks_tolerance_hits():
    """A ray passing 4e-7 outside a triangle's edge: Moller-Trumbore's 1e-6
    barycentric tolerance proposes the triangle, the watertight re-check
    rejects it, and the binary fallback reports the miss (a mode that kept
    the unverified proposal, with its own hit point, would report t = 1).  The second triangle widens
    the wide node, so the first one's quantized box (rounded outward by one
    step of 1/65535 of the parent's extent) contains the ray."""
    pos = np.array([0, 0, 1, 0, 1, 0, 1, 0, 0, 1, 1, 0,
                    2, -1, 1, 0, 3, -1, 1, 0, 2, -0.5, 1, 0], np.float32)
    idx = np.array([0, 1, 2, 0, 3, 4, 5, 0], np.uint32)
    ps = single_instance_scene(pos, idx)
    cw, roots = O.scene_cwbvh(ps)
    r = rays_of([[0.5, -4e-7, 0.0], [0.5, 4e-7, 0.0]], [[0, 0, 1], [0, 0, 1]])
    h = O.trace_closest_cwbvh(ps, cw, roots, r)
    assert h["inst"][0] == 0xFFFFFFFF
    assert h["inst"][1] == 0 and h["t"][1] == np.float32(1.0)
    assert_records_equal(h, O.trace_closest(ps, r))
