"""Device refits against the oracle's restatements (oracle/refit.c).

  * vt_refit_bvh_bounds_device (refitBvhBounds.ts:80-179): the boxes of
    built trees after moved (and unmoved, signed-zero) vertices equal the
    oracle's bit for bit, for 3- and 4-word strides and trees from one
    triangle to 300k; every preflight rejection (refitBvhBounds.ts:3-63)
    returns VT_ERR_BUILD and leaves the caller's nodes untouched.
  * vt_refit_primitive (the reference's refit path for moved geometry,
    HybridEnginePrimitiveUpdates.ts:1482-1560): the bound scene's packed
    BLAS and TLAS nodes equal oracle.refit_primitive's; every box slot of the
    device Node64 and Node128 trees holds a refitted box; closest hits in the
    Node128, Node64 and cwbvh-closest modes equal the oracle's traversal of
    the refitted pack; a frame after the refit meets the path-matching
    parity contract against the binary64 integrator on that pack;
    rejections leave the scene unchanged; a recorded frame graph of the
    engine is invalidated.
"""
import time

import numpy as np
import pytest

from oracle import oracle as O
from vitrum_amd import native as N
from vitrum_amd import workloads as W

from render_cases import cpu_frames
from test_gpu_traverse import assert_hits_equal
from test_oracle_refit import soup

pytestmark = pytest.mark.gpu

SEED = 0x9E3779B9


def device_refit(engine, nodes, idx, pos, istride=4, pstride=4):
    nodes = np.ascontiguousarray(nodes, np.uint32)
    idx = np.ascontiguousarray(idx, np.uint32).ravel()
    pos = np.ascontiguousarray(pos, np.float32).ravel()
    bufs = [engine.alloc(max(a.nbytes, 16)) for a in (nodes, idx, pos)]
    try:
        for d, a in zip(bufs, (nodes, idx, pos)):
            if a.nbytes:
                engine.h2d(d, a)
        engine.refit_bvh_bounds_device(bufs[0], nodes.size // 8, bufs[1], idx.size, istride, bufs[2], pos.size,
                                       pstride)
        out = np.zeros_like(nodes)
        engine.d2h(out, bufs[0])
        return out
    finally:
        for d in bufs:
            engine.free(d)


def moved(p4, seed, scale=0.5):
    rng = np.random.default_rng(seed)
    m = p4.copy()
    m[:, :3] += rng.standard_normal((m.shape[0], 3)).astype(np.float32) * np.float32(scale)
    return m


@pytest.mark.parametrize("n,seed,zeros", [(1, 0, False), (100, 1, False), (20000, 2, True), (300000, 3, False)])
def test_device_refit_equals_oracle(engine, n, seed, zeros):
    p4, i4 = soup(n, seed, zeros)
    b = O.build_blas(p4, i4)
    print(f"refit {n} triangles, {b['node_count']} nodes", flush=True)
    for pos in (p4, moved(p4, seed + 7)):
        want = O.refit_bvh_bounds(b["nodes"], b["reordered_indices"], pos)
        got = device_refit(engine, b["nodes"], b["reordered_indices"], pos)
        np.testing.assert_array_equal(got, want)


def test_device_refit_strides(engine):
    p4, i4 = soup(5000, 11, zeros=True)
    b = O.build_blas(p4, i4)
    pos = moved(p4, 12)
    i3 = b["reordered_indices"].reshape(-1, 4)[:, :3].copy()
    p3 = pos[:, :3].copy()
    for idx, ist in ((b["reordered_indices"], 4), (i3, 3)):
        for p, pst in ((pos, 4), (p3, 3)):
            want = O.refit_bvh_bounds(b["nodes"], idx, p, index_stride=ist, position_stride=pst)
            np.testing.assert_array_equal(device_refit(engine, b["nodes"], idx, p, ist, pst), want)


def test_device_refit_rejections_leave_nodes_untouched(engine):
    p4, i4 = soup(64, 13)
    b = O.build_blas(p4, i4)
    nodes, idx = b["nodes"], b["reordered_indices"]
    leaf = next(i for i in range(nodes.size // 8) if nodes[8 * i + 7] >> 16 == 0xFFFF)
    cases = []
    bad = nodes.copy(); bad[7] = 3; cases.append((bad, idx, p4))
    bad = nodes.copy(); bad[6] = nodes.size // 8; cases.append((bad, idx, p4))
    bad = nodes.copy(); bad[6] = 1; cases.append((bad, idx, p4))
    bad = nodes.copy(); bad[8 * leaf + 6] = idx.size // 4; cases.append((bad, idx, p4))
    bi = idx.copy(); bi[4 * nodes[8 * leaf + 6]] = p4.shape[0]; cases.append((nodes, bi, p4))
    bp = p4.copy(); bp[idx[4 * nodes[8 * leaf + 6]], 2] = np.nan; cases.append((nodes, idx, bp))
    extra = np.concatenate([nodes, np.array([0, 0, 0, 0, 0, 0, 0, 0xFFFF0000], np.uint32)])
    cases.append((extra, idx, p4))
    for k, (nd, ix, ps) in enumerate(cases):
        with pytest.raises(O.OracleError):
            O.refit_bvh_bounds(nd, ix, ps)
        d = engine.alloc(nd.nbytes)
        di = engine.alloc(ix.nbytes)
        dp = engine.alloc(ps.nbytes)
        try:
            engine.h2d(d, nd); engine.h2d(di, ix); engine.h2d(dp, ps)
            with pytest.raises(N.VtError) as ei:
                engine.refit_bvh_bounds_device(d, nd.size // 8, di, ix.size, 4, dp, ps.size, 4)
            assert ei.value.status == -3, (k, str(ei.value))
            out = np.zeros_like(nd)
            engine.d2h(out, d)
            np.testing.assert_array_equal(out, nd)
        finally:
            for x in (d, di, dp):
                engine.free(x)
    with pytest.raises(N.VtError) as ei:
        engine.refit_bvh_bounds_device(0, 1, 0, 0, 5, 0, 0, 4)
    assert ei.value.status == -2


# ---- a bound scene ---------------------------------------------------------

def prims_of(sc):
    return [dict(p, material_id=i) for i, p in enumerate(sc["primitives"])]


def displaced(positions, seed, amp=0.03):
    rng = np.random.default_rng(seed)
    p = np.asarray(positions, np.float32).reshape(-1, 3)
    return (p * np.float32(1.1) + rng.standard_normal(p.shape).astype(np.float32) * np.float32(amp)).astype(np.float32)


def instanced_scene():
    sc = W.spheres_scene(instances=24, lights=4)
    base = sc["primitives"][0]
    mats = [np.asarray(p.get("transform", np.eye(4, dtype=np.float32).ravel()), np.float32)
            for p in sc["primitives"][1:9]]
    inst = dict(kind="instanced-mesh", positions=base["positions"], normals=base["normals"],
                indices=base["indices"], material=base["material"], instances=mats)
    return dict(sc, primitives=[inst] + sc["primitives"][9:])


def slot_boxes(words, per_node, kids):
    """(box words, ref) of every slot of a Node64 (kids=2) / Node128 (kids=4) array."""
    w = np.asarray(words, np.uint32).reshape(-1, per_node)
    if kids == 2:
        boxes = np.concatenate([w[:, 0:6], w[:, 6:12]])
        refs = np.concatenate([w[:, 12], w[:, 13]])
    else:
        boxes = np.stack([w[:, a * 4:(a + 1) * 4] for a in range(6)], axis=2).reshape(-1, 6)
        refs = w[:, 24:28].reshape(-1)
    return boxes, refs


def check_trees_hold_refitted_boxes(engine, ref_nodes):
    allowed = {tuple(r) for r in np.asarray(ref_nodes, np.uint32).reshape(-1, 8)[:, :6].tolist()}
    for field, per, kids in ((N.SCENE_BLAS_NODE64, 16, 2), (N.SCENE_BLAS_NODE128, 32, 4)):
        boxes, refs = slot_boxes(engine.scene_field(field), per, kids)
        live = refs != 0x80000000
        stale = [tuple(b) for b in boxes[live].tolist() if tuple(b) not in allowed]
        assert not stale, f"{field}: {len(stale)} slot boxes are not refitted node boxes"


def scene_rays(sc, res=192, seed=3):
    rays, _ = W.primary_rays(sc["camera"], res, res, scene_radius=sc["scene_radius"])
    rng = np.random.default_rng(seed)
    r = np.zeros(40000, W.RAY_DTYPE)
    r["o"] = rng.uniform(-0.8, 0.8, (r.size, 3)) * sc["scene_radius"] * 0.3
    d = rng.normal(size=(r.size, 3))
    r["d"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    r["tmax"] = W.TMAX_INFINITY
    return np.concatenate([rays, r])


@pytest.mark.parametrize("which", ["spheres", "instanced"])
def test_scene_refit_matches_oracle(engine, which):
    sc = W.spheres_scene(instances=48, lights=4) if which == "spheres" else instanced_scene()
    prims = prims_of(sc)
    prim = 3 if which == "spheres" else 0
    engine.set_scene(sc)
    pack = O.pack_scene(prims)
    new = displaced(prims[prim]["positions"], 21)
    t0 = time.perf_counter()
    engine.refit_primitive(prim, new)
    print(f"{which}: refit of primitive {prim} ({new.shape[0]} vertices) {1e3 * (time.perf_counter() - t0):.2f} ms",
          flush=True)
    want, _, _ = O.refit_primitive(pack, prim, new)
    for f in ("positions", "bvh_nodes", "tlas_nodes"):
        np.testing.assert_array_equal(engine.scene_field(f).view(np.uint32), want[f].view(np.uint32), err_msg=f)
    check_trees_hold_refitted_boxes(engine, want["bvh_nodes"])
    rays = scene_rays(sc)
    ps = O.PackedScene(want)
    hits = O.trace_closest(ps, rays)
    assert_hits_equal(engine.trace_closest(rays), hits, f"{which} refit node128")
    engine.set_node_width(2)
    try:
        assert_hits_equal(engine.trace_closest(rays), hits, f"{which} refit node64")
    finally:
        engine.set_node_width(4)
    assert np.mean(hits["inst"] != 0xFFFFFFFF) > 0.05


def test_scene_refit_in_cwbvh_mode():
    sc = W.spheres_scene(instances=24, lights=4)
    prims = prims_of(sc)
    eng = N.Engine(0, bvh_traversal=1)
    try:
        eng.set_scene(sc)
        rays = scene_rays(sc, res=96)
        eng.trace_closest(rays)  # builds the wide forest of the first geometry
        new = displaced(prims[5]["positions"], 5, amp=0.08)
        eng.refit_primitive(5, new)
        want, _, _ = O.refit_primitive(O.pack_scene(prims), 5, new)
        ps = O.PackedScene(want)
        cw, roots = O.scene_cwbvh(ps)
        assert_hits_equal(eng.trace_closest(rays), O.trace_closest_cwbvh(ps, cw, roots, rays), "cwbvh after refit")
    finally:
        eng.close()


def test_noop_scene_refit_keeps_every_buffer(engine):
    sc = W.cornell_scene(8192)
    engine.set_scene(sc)
    before = {f: engine.scene_field(f).copy() for f in ("positions", "bvh_nodes", "tlas_nodes")}
    dev = {f: engine.scene_field(f).copy() for f in (N.SCENE_BLAS_NODE64, N.SCENE_BLAS_NODE128, N.SCENE_TRIANGLES)}
    pack = O.pack_scene(prims_of(sc))
    for prim in (0, 6, 7):
        pos = np.asarray(sc["primitives"][prim]["positions"], np.float32)
        engine.refit_primitive(prim, pos)
        want, _, _ = O.refit_primitive(pack, prim, pos)
        np.testing.assert_array_equal(want["bvh_nodes"], pack["bvh_nodes"])  # no signed-zero flips here
    for f, v in before.items():
        np.testing.assert_array_equal(engine.scene_field(f).view(np.uint32), v.view(np.uint32), err_msg=f)
    for f, v in dev.items():
        np.testing.assert_array_equal(engine.scene_field(f).view(np.uint32), v.view(np.uint32), err_msg=str(f))


def test_frame_after_refit_matches_fp64_integrator(engine):
    sc = W.cornell_scene(8192)
    prims = prims_of(sc)
    prim = 6  # the tall box, scaled by 1.1 about the origin
    new = displaced(prims[prim]["positions"], 9, amp=0.0)
    res = 64
    engine.set_scene(sc)
    engine.refit_primitive(prim, new)
    ivp = W.inv_view_proj(sc["camera"], res, res)
    engine.render_frame(ivp, res, res, 0, SEED, 8)
    gpu = engine.capture(res, res)[..., :3].astype(np.float64)
    gsig = engine.capture_signatures(res, res)
    want, _, _ = O.refit_primitive(O.pack_scene(prims), prim, new)
    bias = engine.render_info()["ray_bias"]
    assert bias == pytest.approx(float(W.ray_bias(sc["scene_radius"])[0]), rel=1e-6)
    cpu, csig = cpu_frames(sc, res, 1, bias, pack=want)
    cpu, csig = cpu[0], csig[0]
    match = gsig == csig
    assert match.mean() >= 0.999, match.mean()
    ok = (np.abs(gpu - cpu) <= 1e-4 * np.abs(cpu) + 1e-7).all(axis=2) | ~match
    assert ok.mean() >= 0.999
    # the moved box is seen: the frame differs from one of the unmoved scene
    base, _ = cpu_frames(sc, res, 1, bias)
    assert not np.array_equal(base[0], cpu)
    # and it is the frame a rebuild of the moved scene renders: another BVH
    # (other triangle ids, so other signatures) over the same surfaces
    moved_sc = dict(sc, primitives=[dict(p, positions=new.ravel()) if i == prim else p
                                    for i, p in enumerate(sc["primitives"])])
    other = N.Engine(0)
    try:
        other.set_scene(moved_sc)
        other.render_frame(ivp, res, res, 0, SEED, 8)
        rebuilt = other.capture(res, res)[..., :3].astype(np.float64)
    finally:
        other.close()
    close = (np.abs(rebuilt - gpu) <= 1e-5 * np.abs(gpu) + 1e-8).all(axis=2)
    assert close.mean() >= 0.999, close.mean()


def test_scene_refit_rejections_keep_the_scene(engine):
    sc = W.cornell_scene(2048)
    engine.set_scene(sc)
    before = {f: engine.scene_field(f).copy() for f in ("positions", "bvh_nodes", "tlas_nodes")}
    pos = np.asarray(sc["primitives"][6]["positions"], np.float32)
    light = np.asarray(sc["primitives"][5]["positions"], np.float32)
    for args, status in (((5, light), -6),                     # emissive: its emitters would move
                         ((6, pos[:-3]), -2),                  # another vertex count
                         ((6, np.where(np.arange(pos.size) == 4, np.nan, pos)), -2),
                         ((6, pos, np.ones(pos.size - 3, np.float32)), -2),
                         ((99, pos), -2)):                     # no such primitive
        with pytest.raises(N.VtError) as ei:
            engine.refit_primitive(*args)
        assert ei.value.status == status, str(ei.value)
    for f, v in before.items():
        np.testing.assert_array_equal(engine.scene_field(f).view(np.uint32), v.view(np.uint32), err_msg=f)
    fresh = N.Engine(0)
    try:
        with pytest.raises(N.VtError) as ei:
            fresh.refit_primitive(0, pos)
        assert ei.value.status == -6
    finally:
        fresh.close()


def test_refit_invalidates_a_recorded_frame_graph():
    sc = W.cornell_scene(2048)
    eng = N.Engine(0)
    b = None
    try:
        eng.set_scene(sc)
        b = N.RenderBatch(eng)
        b.add(eng)
        fr = {"inv_view_proj": W.inv_view_proj(sc["camera"], 32, 32), "width": 32, "height": 32,
              "frame_index": 0, "frame_seed": SEED, "max_bounces": 4}
        b.capture([fr])
        b.replay(1)
        eng.refit_primitive(6, displaced(sc["primitives"][6]["positions"], 2, amp=0.01))
        with pytest.raises(N.VtError) as ei:
            b.replay(2)
        assert ei.value.status == -6
    finally:
        if b is not None:
            b.close()
        eng.close()


def test_large_blas_refit():
    """The ~200k-triangle floor of a 1M-triangle scene: the packed nodes after
    the device refit equal the oracle's refit of the product's own pack (the
    floor's BLAS is the forest's first tree); later trees are untouched."""
    sc = W.cornell_scene(1_000_000)
    eng = N.Engine(0, build_threads=16)
    try:
        t0 = time.perf_counter()
        eng.set_scene(sc)
        t_set = time.perf_counter() - t0
        nodes = eng.scene_field("bvh_nodes").copy()
        idx = eng.scene_field("indices").copy()
        pos = eng.scene_field("positions").copy().reshape(-1, 4)
        floor = np.asarray(sc["primitives"][0]["positions"], np.float32).reshape(-1, 3)
        new = floor.copy()
        new[:, 1] += np.float32(0.02) * np.sin(np.float32(7.0) * floor[:, 0]) * np.cos(np.float32(5.0) * floor[:, 2])
        eng.refit_primitive(0, new)  # the first refit of a scene uploads its refit links
        new[:, 1] += np.float32(0.01)
        t0 = time.perf_counter()
        eng.refit_primitive(0, new)
        t_refit = time.perf_counter() - t0
        got = eng.scene_field("bvh_nodes")
        pos[: floor.shape[0], :3] = new
        end = 8 * _blas_end(nodes)
        want = O.refit_bvh_bounds(nodes[:end], idx, pos)
        np.testing.assert_array_equal(got[:end], want)
        np.testing.assert_array_equal(got[end:], nodes[end:])
        print(f"1M-triangle scene: set_scene {t_set * 1e3:.0f} ms; refit of the "
              f"{sc['primitives'][0]['indices'].size // 3}-triangle floor {t_refit * 1e3:.2f} ms", flush=True)
    finally:
        eng.close()


def _blas_end(nodes):
    """Node count of the tree rooted at node 0 of a preorder forest."""
    n = np.asarray(nodes, np.uint32).reshape(-1, 8)
    end = 1
    stack = [0]
    while stack:
        i = stack.pop()
        end = max(end, i + 1)
        if n[i, 7] >> 16 != 0xFFFF:
            stack += [i + int(n[i, 6]), i + 1]
    return end
