"""The oracle's builders pinned against the reference's golden vectors and
known-answer tests, and the product's host builders checked against it.

Golden data: tests/golden/scenePack.golden.json is the fixture of the
reference's packages/shared-bvh/src/__tests__/scenePack.golden.json (MIT,
jsquire4/vitrum); the scenes are restated from scenePackGolden.test.ts:45-80
and :226-291.  KATs restate buildArrayBvh.test.ts:44-285 and tlas.test.ts.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from vitrum_amd import native as N
from vitrum_amd import workloads as W

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "scenePack.golden.json")


def box(mn, mx, transform=None):
    x0, y0, z0 = mn
    x1, y1, z1 = mx
    pos = np.array([x0, y0, z0, x1, y0, z0, x0, y1, z0, x0, y0, z1,
                    x1, y1, z1, x1, y0, z1, x0, y1, z1, x1, y1, z1], np.float32)
    nrm = np.array([1 if i % 3 == 2 else 0 for i in range(24)], np.float32)
    idx = np.array([0, 1, 2, 4, 1, 2, 1, 5, 6, 5, 4, 6, 0, 2, 3, 2, 6, 7,
                    0, 1, 3, 1, 5, 3, 3, 5, 7, 5, 6, 7, 0, 4, 3, 4, 6, 7], np.uint32)
    d = dict(kind="mesh", positions=pos, normals=nrm, indices=idx, material_id=0)
    if transform is not None:
        d["transform"] = np.array(transform, np.float32)
    return d


def tri_instances():
    return dict(kind="instanced-mesh", positions=np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float32),
                normals=np.array([0, 0, 1] * 3, np.float32), material_id=0,
                instances=[np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 5, 0, 0, 1], np.float32),
                           np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 7, 0, 0, 1], np.float32)])


TX3 = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 3, 0, 0, 1]
GOLDEN_SCENES = {
    "multiMesh": lambda: [box([0, 0, 0], [1, 1, 1]), box([0, 0, 0], [1, 1, 1], TX3), tri_instances(),
                          box([10, 0, 0], [11, 1, 1])],
    # The splice goldens equal a full pack of the post-splice scene.  Material
    # ids follow the golden's resolver, id.charCodeAt(0) % 7 ('shape-a' -> 3).
    "resizeSplice": lambda: [dict(box([0, 0, 0], [1, 1, 1]), material_id=3),
                             box([5, 0, 0], [6, 1, 1])],
    "sameSizeSplice": lambda: [box([0, 0, 0], [1, 1, 1]), box([0.1, 0.1, 0.1], [1.1, 1.1, 1.1], TX3),
                               tri_instances(), box([10, 0, 0], [11, 1, 1])],
}
FIELD_MAP = [("positions", "positions", True), ("normals", "normals", True),
             ("indices", "indices", False), ("tri_material", "triMaterialIds", False),
             ("bvh_nodes", "bvhNodes", False), ("tlas_nodes", "tlasNodes", False),
             ("tlas_instance_indices", "tlasInstanceIndices", False),
             ("tlas_blas_roots", "tlasBlasRoots", False),
             ("tlas_w2l", "tlasInstanceWorldToLocal", True),
             ("tlas_l2w", "tlasInstanceLocalToWorld", True)]


def bits(a, is_float):
    return np.asarray(a, np.float32).view(np.uint32) if is_float else np.asarray(a, np.uint32)


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


@pytest.mark.parametrize("name", list(GOLDEN_SCENES))
@pytest.mark.parametrize("packer", ["oracle", "product"])
def test_pack_matches_golden_bytes(golden, name, packer):
    prims = GOLDEN_SCENES[name]()
    pack = O.pack_scene(prims) if packer == "oracle" else N.pack_scene(prims)
    g = golden[name]
    for mine, theirs, is_float in FIELD_MAP:
        np.testing.assert_array_equal(bits(pack[mine], is_float), bits(g[theirs], is_float),
                                      err_msg=f"{name}.{theirs}")
    assert pack["tlas_nodes"].size // 8 == g["tlasNodeCount"]
    assert pack["tri_material"].size == g["triangleCount"]


def test_golden_bindings(golden):
    pack = O.pack_scene(GOLDEN_SCENES["multiMesh"]())
    b = pack["bindings"].reshape(-1, 8)
    exp = [(x["blasRoot"], x["instanceCount"], x["vertexStart"], x["vertexCount"], x["triStart"],
            x["triCount"]) for x in golden["multiMesh"]["primitiveTlasBindings"]]
    assert [tuple(r[:6]) for r in b.tolist()] == exp


# ---- buildArrayBvh KATs (buildArrayBvh.test.ts) ----------------------------

def stride4_positions(n):
    i = np.arange(n)
    return np.stack([i, (i % 3) * 0.25, (i % 5) * 0.5, np.zeros(n)], 1).astype(np.float32).ravel()


SIX_TRIS = np.array([0, 1, 2, 0, 3, 4, 5, 0, 6, 7, 8, 0, 9, 10, 11, 0, 12, 13, 14, 0, 15, 16, 17, 0],
                    np.uint32)
BUILDERS = {"oracle": O.build_blas, "product": N.build_blas}


@pytest.mark.parametrize("builder", list(BUILDERS))
def test_blas_empty_is_single_zero_count_leaf(builder):
    r = BUILDERS[builder](np.zeros(0, np.float32), np.zeros(0, np.uint32))
    assert r["node_count"] == 1
    assert r["nodes"][6] == 0 and r["nodes"][7] == 0xFFFF0000


@pytest.mark.parametrize("builder", list(BUILDERS))
def test_blas_single_triangle_leaf(builder):
    r = BUILDERS[builder](stride4_positions(3), np.array([0, 1, 2, 0], np.uint32), np.array([7], np.uint32))
    assert r["node_count"] == 1 and r["nodes"][6] == 0 and r["nodes"][7] == 0xFFFF0001
    assert r["reordered_indices"].tolist() == [0, 1, 2, 0] and r["reordered_mat"].tolist() == [7]


@pytest.mark.parametrize("builder", list(BUILDERS))
def test_blas_multiset_and_offsets(builder):
    r = BUILDERS[builder](stride4_positions(18), SIX_TRIS, np.arange(6, dtype=np.uint32))
    assert r["node_count"] > 1
    assert sorted(r["reordered_mat"].tolist()) == list(range(6))
    assert sorted(r["reordered_to_source"].tolist()) == list(range(6))
    O.validate_bvh(r["nodes"], r["node_count"], triangle_count=6)
    w = r["nodes"].reshape(-1, 8)
    interior = [(i, row) for i, row in enumerate(w) if (row[7] >> 16) != 0xFFFF]
    assert interior and all(1 <= row[6] < len(w) for _, row in interior)


@pytest.mark.parametrize("builder", list(BUILDERS))
def test_blas_stride3(builder):
    pos = stride4_positions(6).reshape(-1, 4)[:, :3].ravel()
    r = BUILDERS[builder](pos, np.arange(6, dtype=np.uint32), np.array([11, 22], np.uint32),
                          position_stride=3, index_stride=3)
    assert r["node_count"] == 1 and r["reordered_indices"].tolist() == list(range(6))
    assert r["reordered_mat"].tolist() == [11, 22]


def coplanar_floor(nquads=1000):
    side = int(np.ceil(np.sqrt(nquads)))
    pos, idx, v, q = [], [], 0, 0
    for i in range(side):
        for j in range(side):
            if q >= nquads:
                break
            pos += [i, 0, j, 0, i + 1, 0, j, 0, i + 1, 0, j + 1, 0, i, 0, j + 1, 0]
            idx += [v, v + 1, v + 2, 0, v, v + 2, v + 3, 0]
            v += 4
            q += 1
    return np.array(pos, np.float32), np.array(idx, np.uint32)


def tree_depth(words):
    w = np.asarray(words).reshape(-1, 8)
    best, leaf_tris, max_leaf, stack = 0, 0, 0, [(0, 0)]
    while stack:
        i, d = stack.pop()
        best = max(best, d)
        if (w[i, 7] >> 16) == 0xFFFF:
            leaf_tris += int(w[i, 7] & 0xFFFF)
            max_leaf = max(max_leaf, int(w[i, 7] & 0xFFFF))
        else:
            stack += [(i + 1, d + 1), (i + int(w[i, 6]), d + 1)]
    return best, leaf_tris, max_leaf


@pytest.mark.parametrize("builder", list(BUILDERS))
def test_blas_coplanar_floor_is_balanced(builder):
    pos, idx = coplanar_floor()
    r = BUILDERS[builder](pos, idx)
    depth, leaf_tris, max_leaf = tree_depth(r["nodes"])
    assert leaf_tris == 2000 and max_leaf <= 4 and depth <= 16


@pytest.mark.parametrize("builder", list(BUILDERS))
def test_blas_coincident_centroids_median_split(builder):
    pos = np.array([0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0], np.float32)
    idx = np.tile(np.array([0, 1, 2, 0], np.uint32), 17)
    r = BUILDERS[builder](pos, idx, max_leaf=3)
    _, leaf_tris, max_leaf = tree_depth(r["nodes"])
    assert leaf_tris == 17 and max_leaf <= 3
    assert r["reordered_to_source"].tolist() == list(range(17))


@pytest.mark.parametrize("builder", list(BUILDERS))
def test_blas_depth_cap_errors(builder):
    err = O.OracleError if builder == "oracle" else N.VtError
    idx = np.array([v for t in range(8) for v in (3 * t, 3 * t + 1, 3 * t + 2, 0)], np.uint32)
    with pytest.raises(err, match="maxDepth=1 prevents enforcing maxLeafTriangles=1"):
        BUILDERS[builder](stride4_positions(24), idx, max_leaf=1, max_depth=1)
    with pytest.raises(err):
        BUILDERS[builder](stride4_positions(3), np.array([0, 1, 2, 0], np.uint32), max_depth=59)
    big = np.tile(np.array([0, 1, 2, 0], np.uint32), 0x10000)
    with pytest.raises(err, match=r"maxDepth=0 .* and the packed 16-bit limit \(0xFFFF = 65535\)"):
        BUILDERS[builder](np.array([0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0], np.float32), big,
                          max_leaf=1, max_depth=0)


@pytest.mark.parametrize("builder", list(BUILDERS))
def test_blas_rejects_bad_input(builder):
    err = O.OracleError if builder == "oracle" else N.VtError
    with pytest.raises(err):
        BUILDERS[builder](np.array([0, 0, 0, 0, np.nan, 0, 0, 0, 0, 1, 0, 0], np.float32),
                          np.array([0, 1, 2, 0], np.uint32))
    with pytest.raises(err):
        BUILDERS[builder](stride4_positions(3), np.array([0, 1, 5, 0], np.uint32))


@pytest.mark.parametrize("builder", list(BUILDERS))
def test_blas_deterministic(builder):
    a = BUILDERS[builder](stride4_positions(18), SIX_TRIS)
    b = BUILDERS[builder](stride4_positions(18), SIX_TRIS)
    assert np.array_equal(a["nodes"], b["nodes"])


# ---- buildTlas KATs (tlas.test.ts:213-250) ---------------------------------

TLAS_BUILDERS = {"oracle": O.build_tlas, "product": N.build_tlas}
IDENT = np.eye(4, dtype=np.float32).ravel()


@pytest.mark.parametrize("builder", list(TLAS_BUILDERS))
def test_tlas_coincident_centroids_balanced(builder):
    n = 2048
    t = TLAS_BUILDERS[builder](np.arange(n), np.zeros((n, 3)), np.ones((n, 3)), np.tile(IDENT, n))
    depth, _, max_leaf = tree_depth(t["nodes"])
    assert depth == 11 and max_leaf == 1
    if builder == "oracle":
        assert t["fallbacks"][0] > 0  # degenerate-centroids
    t2 = TLAS_BUILDERS[builder](np.arange(n), np.zeros((n, 3)), np.ones((n, 3)), np.tile(IDENT, n))
    assert np.array_equal(t["nodes"], t2["nodes"])


@pytest.mark.parametrize("builder", list(TLAS_BUILDERS))
def test_tlas_depth_safety_fallback(builder):
    x = np.array([1_000_000.0 if i == 15 else float(i) for i in range(16)])
    mn = np.stack([x, np.zeros(16), np.zeros(16)], 1)
    mx = np.stack([x + 0.5, np.ones(16), np.ones(16)], 1)
    t = TLAS_BUILDERS[builder](np.arange(16), mn, mx, np.tile(IDENT, 16), max_depth=4)
    depth, _, _ = tree_depth(t["nodes"])
    assert depth == 4 and len(set(t["instance_indices"].tolist())) == 16
    if builder == "oracle":
        assert t["fallbacks"][3] > 0  # depth-safety


@pytest.mark.parametrize("builder", list(TLAS_BUILDERS))
def test_tlas_single_and_pair(builder):
    t = TLAS_BUILDERS[builder]([7], [[0, 0, 0]], [[1, 2, 3]], IDENT)
    assert t["node_count"] == 1 and t["blas_roots"].tolist() == [7]
    assert np.array_equal(t["nodes"][:6].view(np.float32), np.array([0, 0, 0, 1, 2, 3], np.float32))
    t = TLAS_BUILDERS[builder]([1, 2], [[0, 0, 0], [5, 0, 0]], [[1, 1, 1], [6, 2, 1]], np.tile(IDENT, 2))
    assert np.array_equal(t["nodes"][:6].view(np.float32), np.array([0, 0, 0, 6, 2, 1], np.float32))


# ---- invertMat4 ------------------------------------------------------------

def test_invert_mat4_closed_forms():
    m = np.array([2, 0, 0, 0, 0, 4, 0, 0, 0, 0, 8, 0, 1, 2, 3, 1], np.float32)
    inv = O.invert_mat4(m)
    expect = np.array([0.5, 0, 0, 0, 0, 0.25, 0, 0, 0, 0, 0.125, 0, -0.5, -0.5, -0.375, 1], np.float32)
    np.testing.assert_array_equal(inv, expect)
    assert O.invert_mat4(np.zeros(16, np.float32)) is None
    rng = np.random.default_rng(1)
    for _ in range(20):
        a = rng.normal(size=(4, 4)) + 3 * np.eye(4)
        inv = O.invert_mat4(a.T.astype(np.float32).ravel())
        prod = a.T.astype(np.float32).astype(np.float64).reshape(4, 4).T @ inv.reshape(4, 4).T.astype(np.float64)
        np.testing.assert_allclose(prod, np.eye(4), atol=1e-4)


# ---- product vs oracle at config scale -------------------------------------

@pytest.mark.parametrize("name", ["cornell", "spheres"])
def test_product_pack_equals_oracle_on_config_scenes(name):
    sc = W.config_scene(name)
    prims = [dict(p, material_id=i) for i, p in enumerate(sc["primitives"])]
    po = O.pack_scene(prims)
    pp = N.pack_scene(prims, threads=8)
    for k in N.PACK_FIELDS:
        assert np.array_equal(po[k].view(np.uint32), pp[k].view(np.uint32)), k
