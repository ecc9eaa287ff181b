"""CPU pins of the refit restatements (oracle/refit.c).

  * refitBvhBounds (refitBvhBounds.ts:80-179), after the reference's own
    tests (shared-bvh/src/__tests__/refitBvhBounds.test.ts): a refit with
    unchanged positions is byte-identical -- here on the reference's golden
    packed nodes (tests/golden/scenePack.golden.json, whose boxes include
    signed zeros) as well as on built trees; after moving vertices every leaf
    box is its triangles' bounds, every interior box the union of its
    children, the root the brute-force bounds; single-triangle, empty and
    zero-count-leaf trees; stride-4 positions ignore the fourth lane; the
    preflight's rejections (refitBvhBounds.ts:3-63).
  * refitTlasInstances (tlas.ts:1141-1209): refitting with the boxes the TLAS
    was built from reproduces the built nodes byte for byte; a moved instance
    rewrites exactly its leaf and ancestors, children first; rejections.
  * refit_primitive on a pack: unchanged positions leave the pack
    byte-identical; moved positions give the BLAS and TLAS roots the bounds a
    fresh pack of the moved scene has.
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle as O
from vitrum_amd import workloads as W

from test_oracle_build import GOLDEN, GOLDEN_SCENES

LEAF = 0xFFFF0000


def soup(n, seed, zeros=False):
    rng = np.random.default_rng(seed)
    pos = rng.standard_normal((3 * n, 3)).astype(np.float32)
    if zeros:
        pos[rng.random(pos.shape) < 0.2] = 0.0
        pos[rng.random(pos.shape) < 0.1] = -0.0
    p4 = np.concatenate([pos, np.zeros((3 * n, 1), np.float32)], 1)
    i4 = np.concatenate([np.arange(3 * n, dtype=np.uint32).reshape(-1, 3), np.zeros((n, 1), np.uint32)], 1)
    return p4, i4


def nodes_f(nodes):
    return np.asarray(nodes, np.uint32).reshape(-1, 8)[:, :6].view(np.float32)


def brute_leaf(nodes, i, idx, pos, istride=4, pstride=4):
    q = np.asarray(nodes, np.uint32).reshape(-1, 8)[i]
    off, cnt = int(q[6]), int(q[7]) & 0xFFFF
    tri = np.asarray(idx, np.uint32).reshape(-1, istride)[off:off + cnt, :3].ravel()
    v = np.asarray(pos, np.float32).reshape(-1, pstride)[tri, :3]
    return np.concatenate([v.min(0), v.max(0)])


def check_tree(nodes, idx, pos, istride=4, pstride=4):
    """Every leaf bounds its triangles, every interior node its children."""
    n = np.asarray(nodes, np.uint32).reshape(-1, 8)
    f = nodes_f(nodes)
    for i in range(n.shape[0]):
        if n[i, 7] >> 16 == 0xFFFF:
            if n[i, 7] & 0xFFFF:
                np.testing.assert_array_equal(f[i], brute_leaf(nodes, i, idx, pos, istride, pstride))
        else:
            l, r = i + 1, i + int(n[i, 6])
            np.testing.assert_array_equal(f[i, :3], np.minimum(f[l, :3], f[r, :3]))
            np.testing.assert_array_equal(f[i, 3:], np.maximum(f[l, 3:], f[r, 3:]))


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as fh:
        return json.load(fh)


@pytest.mark.parametrize("name", list(GOLDEN_SCENES))
def test_noop_refit_of_golden_nodes_is_byte_identical(golden, name):
    g = golden[name]
    nodes = np.asarray(g["bvhNodes"], np.uint32)
    pos = np.asarray(g["positions"], np.float32)
    idx = np.asarray(g["indices"], np.uint32)
    roots = sorted(set(int(b["blasRoot"]) for b in g["primitiveTlasBindings"]))
    total = nodes.size // 8
    for k, root in enumerate(roots):
        end = roots[k + 1] if k + 1 < len(roots) else total
        sub = nodes[8 * root:8 * end]
        np.testing.assert_array_equal(O.refit_bvh_bounds(sub, idx, pos), sub)


@pytest.mark.parametrize("n,seed", [(1, 0), (7, 1), (5000, 3)])
def test_noop_refit_of_built_tree_is_byte_identical(n, seed):
    p4, i4 = soup(n, seed)
    b = O.build_blas(p4, i4)
    np.testing.assert_array_equal(O.refit_bvh_bounds(b["nodes"], b["reordered_indices"], p4), b["nodes"])


def test_noop_refit_differs_from_the_build_only_in_zero_signs():
    """The builder takes bounds with Math.min / Math.max (buildArrayBvh.ts:
    171-185: -0 wins a minimum, +0 a maximum), the refit with strict
    comparisons in visiting order (refitBvhBounds.ts:140-170: the first of
    equal values stays).  With mixed signed zeros a no-op refit therefore
    changes only the sign of some zero coordinates."""
    p4, i4 = soup(3000, 2, zeros=True)
    b = O.build_blas(p4, i4)
    r = O.refit_bvh_bounds(b["nodes"], b["reordered_indices"], p4)
    diff = r != b["nodes"]
    assert diff.any()
    assert set(np.unique(r[diff] ^ b["nodes"][diff]).tolist()) == {0x80000000}
    np.testing.assert_array_equal(nodes_f(r), nodes_f(b["nodes"]))  # equal as numbers


@pytest.mark.parametrize("n,seed", [(1, 4), (64, 5), (3000, 6)])
def test_moved_vertices_bound_exactly(n, seed):
    p4, i4 = soup(n, seed)
    b = O.build_blas(p4, i4)
    rng = np.random.default_rng(seed + 100)
    moved = p4.copy()
    moved[:, :3] += rng.standard_normal((moved.shape[0], 3)).astype(np.float32) * np.float32(0.5)
    r = O.refit_bvh_bounds(b["nodes"], b["reordered_indices"], moved)
    assert not np.array_equal(r, b["nodes"])
    np.testing.assert_array_equal(r.reshape(-1, 8)[:, 6:], b["nodes"].reshape(-1, 8)[:, 6:])  # topology kept
    check_tree(r, b["reordered_indices"], moved)
    root = nodes_f(r)[0]
    np.testing.assert_array_equal(root, np.concatenate([moved[:, :3].min(0), moved[:, :3].max(0)]))


def test_single_triangle_leaf_is_exact():
    p4, i4 = soup(1, 9)
    b = O.build_blas(p4, i4)
    moved = np.array([[10, 20, 30, 0], [11, 20, 30, 0], [10, 21, 30, 0]], np.float32)
    r = O.refit_bvh_bounds(b["nodes"], b["reordered_indices"], moved)
    np.testing.assert_array_equal(nodes_f(r)[0], [10, 20, 30, 11, 21, 30])


def test_empty_and_zero_count_leaf():
    assert O.refit_bvh_bounds(np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.float32)).size == 0
    canon = np.array([0, 0, 0, 0, 0, 0, 0, LEAF], np.uint32)
    canon[:3] = np.array([5, 5, 5], np.float32).view(np.uint32)
    r = O.refit_bvh_bounds(canon, np.zeros(0, np.uint32), np.zeros(0, np.float32))
    np.testing.assert_array_equal(r, [0, 0, 0, 0, 0, 0, 0, LEAF])


def test_stride3_indices_and_stride4_positions_ignore_w():
    p4, i4 = soup(200, 11)
    b = O.build_blas(p4, i4)
    i3 = b["reordered_indices"].reshape(-1, 4)[:, :3].copy()
    moved = p4.copy()
    moved[:, :3] *= np.float32(2.0)
    moved[:, 3] = np.nan  # the fourth lane is not a coordinate
    r = O.refit_bvh_bounds(b["nodes"], i3, moved, index_stride=3)
    check_tree(r, i3, moved, istride=3)
    p3 = moved[:, :3].copy()
    np.testing.assert_array_equal(O.refit_bvh_bounds(b["nodes"], i3, p3, position_stride=3, index_stride=3), r)


def _two_leaf_tree():
    p4, i4 = soup(8, 12)
    b = O.build_blas(p4, i4, max_leaf=4)
    assert b["node_count"] >= 3
    return b, p4


def test_preflight_rejections():
    b, p4 = _two_leaf_tree()
    nodes, idx = b["nodes"], b["reordered_indices"]
    with pytest.raises(O.RangeError):
        O.refit_bvh_bounds(nodes, idx, p4, position_stride=5)
    with pytest.raises(O.RangeError):
        O.refit_bvh_bounds(nodes, idx, p4, index_stride=2)
    with pytest.raises(O.OracleError, match="stride-aligned"):
        O.refit_bvh_bounds(nodes, idx[:-1], p4)
    bad = nodes.copy(); bad[7] = 3  # split word of an interior node
    with pytest.raises(O.OracleError, match="interior node 0 is corrupt"):
        O.refit_bvh_bounds(bad, idx, p4)
    bad = nodes.copy(); bad[6] = nodes.size // 8  # right child past the end
    with pytest.raises(O.OracleError, match="corrupt"):
        O.refit_bvh_bounds(bad, idx, p4)
    bad = nodes.copy(); bad[6] = 1  # right child == left child
    with pytest.raises(O.OracleError, match="repeated"):
        O.refit_bvh_bounds(bad, idx, p4)
    leaf = next(i for i in range(nodes.size // 8) if nodes[8 * i + 7] >> 16 == 0xFFFF)
    bad = nodes.copy(); bad[8 * leaf + 6] = idx.size // 4
    with pytest.raises(O.OracleError, match="triangle range"):
        O.refit_bvh_bounds(bad, idx, p4)
    bad_idx = idx.copy(); bad_idx[4 * nodes[8 * leaf + 6]] = p4.shape[0]
    with pytest.raises(O.OracleError, match="invalid vertex"):
        O.refit_bvh_bounds(nodes, bad_idx, p4)
    bad_pos = p4.copy(); bad_pos[idx[4 * nodes[8 * leaf + 6]], 1] = np.inf
    with pytest.raises(O.OracleError, match="non-finite"):
        O.refit_bvh_bounds(nodes, idx, bad_pos)
    extra = np.concatenate([nodes, np.array([0, 0, 0, 0, 0, 0, 0, LEAF], np.uint32)])
    with pytest.raises(O.OracleError, match="unreachable"):
        O.refit_bvh_bounds(extra, idx, p4)


# ---- refitTlasInstances ---------------------------------------------------

def _instance_boxes(pack):
    b = pack["bindings"].reshape(-1, 8)
    pos = pack["positions"].reshape(-1, 4)
    local = {int(r[6]): O.local_aabb(pos[int(r[2]):int(r[2]) + int(r[3])]) for r in b}
    l2w = pack["tlas_l2w"].reshape(-1, 16)
    boxes = [O.transform_aabb(l2w[i], *local[int(p)]) for i, p in enumerate(pack["instance_primitive"])]
    return np.array([m for m, _ in boxes]), np.array([x for _, x in boxes])


@pytest.mark.parametrize("name", list(GOLDEN_SCENES))
def test_tlas_refit_with_build_boxes_reproduces_golden(golden, name):
    pack = O.pack_scene(GOLDEN_SCENES[name]())
    mn, mx = _instance_boxes(pack)
    n = mn.shape[0]
    got, affected = O.refit_tlas_instances(pack["tlas_nodes"], pack["tlas_instance_indices"], mn, mx, np.arange(n))
    np.testing.assert_array_equal(got, np.asarray(golden[name]["tlasNodes"], np.uint32))
    assert sorted(affected.tolist()) == list(range(pack["tlas_nodes"].size // 8))
    assert list(affected) == sorted(affected, reverse=True)  # children before parents


def test_tlas_refit_moves_exactly_one_leaf_and_its_ancestors():
    sc = W.spheres_scene(instances=48, lights=4)
    pack = O.pack_scene([dict(p, material_id=i) for i, p in enumerate(sc["primitives"])])
    mn, mx = _instance_boxes(pack)
    k = 17
    mn2, mx2 = mn.copy(), mx.copy()
    mn2[k] -= 3.0
    mx2[k] += 1.25
    old = pack["tlas_nodes"].reshape(-1, 8)
    got, affected = O.refit_tlas_instances(pack["tlas_nodes"], pack["tlas_instance_indices"], mn2, mx2, [k])
    g = got.reshape(-1, 8)
    # ancestors of k's leaf, from the topology alone
    perm = pack["tlas_instance_indices"]
    parent = {}
    for i, q in enumerate(old):
        if q[7] >> 16 != 0xFFFF:
            parent[i + 1] = i
            parent[i + int(q[6])] = i
    leaf = next(i for i, q in enumerate(old) if q[7] >> 16 == 0xFFFF and perm[q[6]] == k)
    chain = [leaf]
    while chain[-1] in parent:
        chain.append(parent[chain[-1]])
    assert list(affected) == sorted(chain, reverse=True)
    untouched = np.setdiff1d(np.arange(old.shape[0]), chain)
    np.testing.assert_array_equal(g[untouched], old[untouched])
    gf = g[:, :6].view(np.float32)
    np.testing.assert_array_equal(gf[leaf], np.concatenate([mn2[k], mx2[k]]).astype(np.float32))
    for i in chain[1:]:
        l, r = i + 1, i + int(old[i, 6])
        np.testing.assert_array_equal(gf[i, :3], np.minimum(gf[l, :3], gf[r, :3]))
        np.testing.assert_array_equal(gf[i, 3:], np.maximum(gf[l, 3:], gf[r, 3:]))


def test_tlas_refit_rejections():
    pack = O.pack_scene(GOLDEN_SCENES["multiMesh"]())
    mn, mx = _instance_boxes(pack)
    bad = mx.copy(); bad[1, 0] = mn[1, 0] - 1
    with pytest.raises(O.OracleError, match="inverted"):
        O.refit_tlas_instances(pack["tlas_nodes"], pack["tlas_instance_indices"], mn, bad, [1])
    bad = mx.copy(); bad[2, 2] = math.inf
    with pytest.raises(O.OracleError, match="non-finite"):
        O.refit_tlas_instances(pack["tlas_nodes"], pack["tlas_instance_indices"], mn, bad, [0])
    with pytest.raises(O.OracleError, match="out of bounds"):
        O.refit_tlas_instances(pack["tlas_nodes"], pack["tlas_instance_indices"], mn, mx, [mn.shape[0]])


# ---- a primitive's refit on a pack ---------------------------------------

def _prims(sc):
    return [dict(p, material_id=i) for i, p in enumerate(sc["primitives"])]


def test_refit_primitive_with_unchanged_positions_keeps_the_pack():
    sc = W.cornell_scene(2048)
    pack = O.pack_scene(_prims(sc))
    for prim in range(len(sc["primitives"])):
        if not np.any(pack["bindings"].reshape(-1, 8)[:, 6] == prim):
            continue
        pos = np.asarray(sc["primitives"][prim]["positions"], np.float32)
        out, _, _ = O.refit_primitive(pack, prim, pos)
        for f in ("positions", "bvh_nodes", "tlas_nodes", "binding_aabb"):
            np.testing.assert_array_equal(out[f].view(np.uint32), pack[f].view(np.uint32), err_msg=f)


def test_refit_primitive_roots_match_a_fresh_pack():
    sc = W.spheres_scene(instances=24, lights=4)
    prims = _prims(sc)
    pack = O.pack_scene(prims)
    prim = 0
    rng = np.random.default_rng(5)
    pos = np.asarray(prims[prim]["positions"], np.float32).reshape(-1, 3)
    moved = (pos * np.float32(1.3) + rng.standard_normal(pos.shape).astype(np.float32) * np.float32(0.05))
    out, wmn, wmx = O.refit_primitive(pack, prim, moved)
    fresh = O.pack_scene([dict(p, positions=moved.ravel()) if i == prim else p for i, p in enumerate(prims)])
    b = out["bindings"].reshape(-1, 8)
    fb = fresh["bindings"].reshape(-1, 8)
    row = int(np.nonzero(b[:, 6] == prim)[0][0])
    root, froot = int(b[row, 0]), int(fb[row, 0])
    np.testing.assert_array_equal(nodes_f(out["bvh_nodes"])[root], nodes_f(fresh["bvh_nodes"])[froot])
    np.testing.assert_array_equal(nodes_f(out["tlas_nodes"])[0], nodes_f(fresh["tlas_nodes"])[0])
    np.testing.assert_array_equal(out["binding_aabb"], fresh["binding_aabb"])
    roots = sorted(int(x) for x in b[:, 0])
    end = next((x for x in roots if x > root), out["bvh_nodes"].size // 8)
    check_tree(out["bvh_nodes"][8 * root:8 * end], out["indices"], out["positions"])
    # the TLAS root is the binary32 rounding of the binary64 world bounds
    np.testing.assert_array_equal(nodes_f(out["tlas_nodes"])[0], np.concatenate([wmn, wmx]).astype(np.float32))
