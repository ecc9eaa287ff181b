"""The device BLAS builder (vt_build_blas_device) against the oracle's
buildArrayBvh restatement (oracle/bvh_build.c, buildArrayBvh.ts:342-780) and
the host builder: nodes, reordered indices, materials and the permutation
equal word for word, for triangle soups, grids full of tied centroids, flat
and degenerate meshes that take the median fallback, 3- and 4-word strides,
every leaf size and bin count the reference tests use, one triangle to 300k;
argument and depth-limit errors carry the host builder's code and message.
"""
import time

import numpy as np
import pytest

from oracle import oracle as O
from vitrum_amd import native as N

from test_oracle_refit import soup

pytestmark = pytest.mark.gpu


def grid(nx, ny, z_jitter=0.0, seed=0):
    """A tessellated plane: two triangles per cell, many equal centroids."""
    rng = np.random.default_rng(seed)
    xs, ys = np.meshgrid(np.arange(nx + 1, dtype=np.float32), np.arange(ny + 1, dtype=np.float32), indexing="ij")
    z = (rng.standard_normal(xs.shape) * z_jitter).astype(np.float32)
    pos = np.stack([xs, ys, z, np.zeros_like(xs)], -1).reshape(-1, 4)
    v = lambda i, j: i * (ny + 1) + j
    tris = []
    for i in range(nx):
        for j in range(ny):
            tris.append((v(i, j), v(i + 1, j), v(i, j + 1), 0))
            tris.append((v(i + 1, j), v(i + 1, j + 1), v(i, j + 1), 0))
    return pos, np.array(tris, np.uint32)


def assert_same(a, b, what):
    for k in ("node_count", "nodes", "reordered_indices", "reordered_mat", "reordered_to_source"):
        x, y = a[k], b[k]
        if k == "node_count":
            assert x == y, (what, k, x, y)
        else:
            assert np.array_equal(np.asarray(x, np.uint32), np.asarray(y, np.uint32)), (what, k)


def cases():
    p, i = soup(1, 1)
    yield "one triangle", p, i, {}
    p, i = soup(257, 2)
    yield "soup 257 leaf1", p, i, {"max_leaf": 1}
    p, i = soup(3000, 3, zeros=True)
    yield "soup 3000 signed zeros", p, i, {}
    p, i = grid(40, 37)
    yield "flat grid", p, i, {"max_leaf": 2}
    p, i = grid(23, 31, 0.25, 4)
    yield "bumpy grid bins 7", p, i, {"num_bins": 7}
    # every triangle at one point: no centroid span, median fallback all the way down
    p = np.zeros((3 * 200, 4), np.float32)
    p[:, :3] = 1.5
    i = np.concatenate([np.arange(600, dtype=np.uint32).reshape(-1, 3), np.zeros((200, 1), np.uint32)], 1)
    yield "coincident", p, i, {}
    # two clusters far apart plus identical copies: SAH then fallback
    p, i = soup(500, 5)
    p[: 3 * 250, :3] *= 1e-7
    yield "tiny cluster", p, i, {"num_bins": 2, "max_leaf": 3}
    p, i = soup(4096, 6)
    yield "soup 4096 bins 256 leaf 16", p, i, {"num_bins": 256, "max_leaf": 16}
    p, i = soup(999, 7)
    yield "stride 3", np.ascontiguousarray(p[:, :3]), np.ascontiguousarray(i[:, :3]), \
        {"position_stride": 3, "index_stride": 3}
    # ranges above 4096 triangles are reduced, binned and partitioned by
    # several workgroups: ties, flat axes and the fallback there too
    p, i = grid(130, 131)
    yield "wide flat grid", p, i, {"max_leaf": 2}
    p = np.zeros((3 * 40000, 4), np.float32)
    p[:, :3] = -2.0
    i = np.concatenate([np.arange(120000, dtype=np.uint32).reshape(-1, 3), np.zeros((40000, 1), np.uint32)], 1)
    yield "wide coincident", p, i, {"max_leaf": 8}
    p, i = soup(70000, 8, zeros=True)
    yield "wide soup bins 32", p, i, {"num_bins": 32}
    p, i = soup(50000, 9)
    p[:25000, :3] = np.round(p[:25000, :3])
    yield "wide rounded soup leaf 1", p, i, {"max_leaf": 1, "num_bins": 5}


@pytest.mark.parametrize("case", list(range(13)))
def test_device_build_matches_oracle(engine, case):
    name, p, i, kw = list(cases())[case]
    n = i.shape[0]
    mats = (np.arange(n, dtype=np.uint32) * 7) % 5
    print("device build case", name, n, kw, flush=True)
    ref = O.build_blas(p, i, mats, **kw)
    got = N.build_blas(p, i, mats, engine=engine, **kw)
    assert_same(got, ref, name)


def test_device_build_large_matches_host(engine):
    p, i = soup(300_000, 11)
    print("device build 300k soup", flush=True)
    t0 = time.perf_counter()
    host = N.build_blas(p, i)
    t1 = time.perf_counter()
    dev = N.build_blas(p, i, engine=engine)
    t2 = time.perf_counter()
    assert_same(dev, host, "300k soup")
    print(f"host build {1e3 * (t1 - t0):.1f} ms, device build {1e3 * (t2 - t1):.1f} ms", flush=True)


def test_device_build_errors_match_host(engine):
    p, i = soup(64, 12)
    for kw in ({"max_leaf": 0}, {"num_bins": 1}, {"max_depth": 59}, {"max_depth": 2}, {"max_depth": 0}):
        with pytest.raises(N.VtError) as host:
            N.build_blas(p, i, **kw)
        with pytest.raises(N.VtError) as dev:
            N.build_blas(p, i, engine=engine, **kw)
        assert str(dev.value) == str(host.value), kw
    bad = i.copy()
    bad[5, 1] = 10_000
    with pytest.raises(N.VtError, match="out-of-range vertex"):
        N.build_blas(p, bad, engine=engine)
    nan = p.copy()
    nan[7, 2] = np.nan
    with pytest.raises(N.VtError, match="must be finite"):
        N.build_blas(nan, i, engine=engine)
    # the engine stays usable after a rejected build
    got = N.build_blas(p, i, engine=engine)
    assert_same(got, N.build_blas(p, i), "after errors")


def test_device_build_empty(engine):
    got = N.build_blas(np.zeros(0, np.float32), np.zeros(0, np.uint32), engine=engine)
    host = N.build_blas(np.zeros(0, np.float32), np.zeros(0, np.uint32))
    assert got["node_count"] == host["node_count"] == 1
    assert np.array_equal(got["nodes"], host["nodes"])


@pytest.mark.parametrize("scene_name,threshold", [("cornell", 2000), ("spheres", 2000), ("cornell", 1),
                                                  ("spheres", 1)])
def test_scene_bound_with_device_builds_is_the_same_pack(scene_name, threshold):
    """vt_set_device_build: meshes at or above the threshold are built
    together on the device (one tree per mesh, grown level by level) when
    the scene is bound; every field of the bound pack and the device trees
    equal the host-built ones."""
    from vitrum_amd import workloads as W
    scene = W.cornell_scene(65536) if scene_name == "cornell" else W.spheres_scene(64)
    host, dev = N.Engine(0), N.Engine(0)
    try:
        dev.set_device_build(threshold)
        print("bind", scene_name, flush=True)
        host.set_scene(scene)
        dev.set_scene(scene)
        for field in N.PACK_FIELDS:
            a, b = host.scene_field(field), dev.scene_field(field)
            assert a.dtype == b.dtype and np.array_equal(a.view(np.uint32), b.view(np.uint32)), field
        for field in (N.SCENE_BLAS_NODE64, N.SCENE_BLAS_NODE128):
            assert np.array_equal(host.scene_field(field), dev.scene_field(field)), field
    finally:
        host.close()
        dev.close()


def test_forest_of_many_meshes_matches_host():
    """Several hundred meshes of different sizes (1 to 5000 triangles, some
    empty or degenerate) bound with every BLAS built on the device: the pack
    equals the host-built one."""
    rng = np.random.default_rng(21)
    prims = []
    for m in range(300):
        n = int(rng.integers(1, 5000)) if m % 37 else 1
        p, i = soup(n, 100 + m)
        if m % 11 == 0:
            p[:, :3] = 0.25  # every centroid equal: median fallback
        pos = np.ascontiguousarray(p[:, :3]).ravel()
        prims.append({"kind": "mesh", "material": {"baseColor": [0.5, 0.5, 0.5]}, "positions": pos,
                      "normals": np.zeros_like(pos), "indices": np.ascontiguousarray(i[:, :3]).ravel()})
    scene = {"primitives": prims}
    host, dev = N.Engine(0), N.Engine(0)
    try:
        dev.set_device_build(1)
        print("bind 300 meshes", flush=True)
        host.set_scene(scene)
        dev.set_scene(scene)
        for field in N.PACK_FIELDS:
            a, b = host.scene_field(field), dev.scene_field(field)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), field
    finally:
        host.close()
        dev.close()


def test_device_build_10m_triangles_matches_host(engine):
    """Ten million triangles in one mesh (the scale of the largest
    benchmark scene's triangle count) build on the device to the host
    builder's output, word for word."""
    rng = np.random.default_rng(13)
    n = 10_000_000
    pos = np.zeros((3 * n, 4), np.float32)
    centers = rng.random((n, 3), dtype=np.float32) * 100.0
    for k in range(3):
        pos[k::3, :3] = centers + rng.standard_normal((n, 3), dtype=np.float32) * 0.05
    idx = np.zeros((n, 4), np.uint32)
    idx[:, :3] = np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)
    print("device build 10M", flush=True)
    t0 = time.perf_counter()
    dev = N.build_blas(pos, idx, engine=engine)
    t1 = time.perf_counter()
    print(f"device build {1e3 * (t1 - t0):.0f} ms, {dev['node_count']} nodes", flush=True)
    host = N.build_blas(pos, idx)
    print(f"host build {1e3 * (time.perf_counter() - t1):.0f} ms", flush=True)
    assert_same(dev, host, "10M")
