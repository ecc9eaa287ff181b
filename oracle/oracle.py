"""ctypes binding of the CPU oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

U32P = C.POINTER(C.c_uint32)
F32P = C.POINTER(C.c_float)
F64P = C.POINTER(C.c_double)


class OracleError(RuntimeError):
    pass


class RangeError(OracleError):
    """A rejection the reference raises as RangeError."""


class Primitive(C.Structure):
    _fields_ = [
        ("kind", C.c_uint32),
        ("material_id", C.c_uint32),
        ("positions", F32P),
        ("position_floats", C.c_uint32),
        ("normals", F32P),
        ("normal_floats", C.c_uint32),
        ("indices", U32P),
        ("index_count", C.c_uint32),
        ("transforms", F32P),
        ("transform_count", C.c_uint32),
    ]


class Scene(C.Structure):
    _fields_ = [
        ("positions", F32P),
        ("vertex_count", C.c_uint32),
        ("indices", U32P),
        ("triangle_count", C.c_uint32),
        ("bvh_nodes", U32P),
        ("bvh_node_count", C.c_uint32),
        ("tlas_nodes", U32P),
        ("tlas_node_count", C.c_uint32),
        ("instance_indices", U32P),
        ("instance_index_count", C.c_uint32),
        ("blas_roots", U32P),
        ("w2l", F32P),
        ("l2w", F32P),
        ("instance_count", C.c_uint32),
    ]


FIELDS = [
    "positions", "normals", "indices", "tri_material", "tri_source", "tri_primitive",
    "bvh_nodes", "tlas_nodes", "tlas_instance_indices", "tlas_blas_roots", "tlas_w2l",
    "tlas_l2w", "instance_primitive", "instance_source", "bindings", "binding_aabb",
]
FLOAT_FIELDS = {"positions", "normals", "tlas_w2l", "tlas_l2w", "binding_aabb"}

RAY_DTYPE = np.dtype([("o", "<f4", 3), ("tmin", "<f4"), ("d", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("prim", "<u4"), ("inst", "<u4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4")])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liborc.so")
        if not os.path.exists(path):
            raise OracleError(f"oracle library not built: {path} (run make -C oracle)")
        L = C.CDLL(path)
        L.orc_last_error.restype = C.c_char_p
        L.orc_pack_field_size.restype = C.c_int64
        L.orc_pack_field_size.argtypes = [C.c_void_p, C.c_int]
        L.orc_pack_copy_field.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_uint64]
        L.orc_pack_free.argtypes = [C.c_void_p]
        L.orc_pack_warning_count.argtypes = [C.c_void_p]
        L.orc_pack_warning_count.restype = C.c_uint32
        L.orc_pcg_init.restype = C.c_uint32
        L.orc_pcg_next.restype = C.c_uint32
        L.orc_rand_f32.restype = C.c_float
        L.orc_pcg_init.argtypes = [C.c_uint32] * 3
        L.orc_pcg_next.argtypes = [U32P]
        L.orc_rand_f32.argtypes = [U32P]
        L.orc_watertight.argtypes = [F32P, F32P, F32P, F32P, F32P, C.c_float, F32P]
        L.orc_validate_bvh.argtypes = [U32P, C.c_uint32, U32P, C.c_int32, C.c_int64, C.c_uint32, U32P]
        L.orc_invert_mat4.argtypes = [F32P, F32P]
        _LIB = L
    return _LIB


def _err():
    return lib().orc_last_error().decode()


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def build_blas(positions, indices, tri_material_ids=None, *, position_stride=4, index_stride=4,
               max_leaf=4, num_bins=16, max_depth=58):
    """buildArrayBvh (packages/shared-bvh/src/buildArrayBvh.ts:342)."""
    positions = np.ascontiguousarray(positions, dtype=np.float32).ravel()
    indices = np.ascontiguousarray(indices, dtype=np.uint32).ravel()
    tri_count = indices.size // index_stride
    mats = (np.zeros(tri_count, np.uint32) if tri_material_ids is None
            else np.ascontiguousarray(tri_material_ids, dtype=np.uint32))
    cap = max(1, 2 * tri_count - 1)
    nodes = np.zeros(cap * 8, np.uint32)
    count = C.c_uint32(0)
    ridx = np.zeros(max(indices.size, 1), np.uint32)
    rmat = np.zeros(max(tri_count, 1), np.uint32)
    rsrc = np.zeros(max(tri_count, 1), np.uint32)
    rc = lib().orc_build_blas(
        _p(positions, F32P), C.c_uint32(positions.size), C.c_uint32(position_stride),
        _p(indices, U32P), C.c_uint32(indices.size), C.c_uint32(index_stride), _p(mats, U32P),
        C.c_uint32(max_leaf), C.c_uint32(num_bins), C.c_uint32(max_depth), _p(nodes, U32P),
        C.byref(count), _p(ridx, U32P), _p(rmat, U32P), _p(rsrc, U32P))
    if rc != 0:
        raise OracleError(_err())
    n = count.value
    return {
        "nodes": nodes[: n * 8].copy(),
        "node_count": n,
        "reordered_indices": ridx[: indices.size].copy(),
        "reordered_mat": rmat[:tri_count].copy(),
        "reordered_to_source": rsrc[:tri_count].copy(),
    }


def validate_bvh(node_words, node_count, roots=None, triangle_count=None, stack_capacity=60):
    node_words = np.ascontiguousarray(node_words, dtype=np.uint32)
    r = None if roots is None else np.ascontiguousarray(roots, dtype=np.uint32)
    proof = np.zeros(7, np.uint32)
    rc = lib().orc_validate_bvh(_p(node_words, U32P), node_count, _p(r, U32P),
                                -1 if r is None else r.size,
                                -1 if triangle_count is None else triangle_count,
                                stack_capacity, _p(proof, U32P))
    if rc != 0:
        raise OracleError(_err())
    return dict(zip(["nodeCount", "rootCount", "interiorNodeCount", "leafNodeCount", "maxDepth",
                     "maxTraversalStackEntries", "traversalStackCapacity"], proof.tolist()))


def build_tlas(blas_ids, aabb_min, aabb_max, w2l, *, max_leaf=1, num_bins=16, max_depth=58):
    """buildTlas (packages/shared-bvh/src/tlas.ts:859)."""
    blas_ids = np.ascontiguousarray(blas_ids, dtype=np.uint32)
    n = blas_ids.size
    mn = np.ascontiguousarray(aabb_min, dtype=np.float64).ravel()
    mx = np.ascontiguousarray(aabb_max, dtype=np.float64).ravel()
    xf = np.ascontiguousarray(w2l, dtype=np.float32).ravel()
    nodes = np.zeros(max(1, 2 * n - 1) * 8, np.uint32)
    count = C.c_uint32(0)
    perm = np.zeros(max(n, 1), np.uint32)
    roots = np.zeros(max(n, 1), np.uint32)
    tf = np.zeros(max(n, 1) * 16, np.float32)
    fb = np.zeros(4, np.uint32)
    rc = lib().orc_build_tlas(
        C.c_uint32(n), _p(blas_ids, U32P), _p(mn, F64P), _p(mx, F64P), _p(xf, F32P),
        C.c_uint32(max_leaf), C.c_uint32(num_bins), C.c_uint32(max_depth), _p(nodes, U32P),
        C.byref(count), _p(perm, U32P), _p(roots, U32P), _p(tf, F32P), _p(fb, U32P))
    if rc != 0:
        raise OracleError(_err())
    c = count.value
    return {"nodes": nodes[: c * 8].copy(), "node_count": c, "instance_indices": perm[:n].copy(),
            "blas_roots": roots[:n].copy(), "transforms": tf[: n * 16].copy(),
            "fallbacks": fb.tolist()}


def invert_mat4(m):
    m = np.ascontiguousarray(m, dtype=np.float32).ravel()
    out = np.zeros(16, np.float32)
    ok = lib().orc_invert_mat4(_p(m, F32P), _p(out, F32P))
    return out if ok else None


def refit_bvh_bounds(nodes, indices, positions, *, position_stride=4, index_stride=4):
    """refitBvhBounds (packages/shared-bvh/src/refitBvhBounds.ts:80-179) of the
    tree rooted at node 0 of `nodes`; returns the refitted copy."""
    out = np.array(nodes, dtype=np.uint32).ravel().copy()
    idx = np.ascontiguousarray(indices, dtype=np.uint32).ravel()
    pos = np.ascontiguousarray(positions, dtype=np.float32).ravel()
    rc = lib().orc_refit_bvh_bounds(_p(out, U32P), C.c_uint32(out.size // 8), _p(idx, U32P),
                                    C.c_uint32(idx.size), C.c_uint32(index_stride), _p(pos, F32P),
                                    C.c_uint32(pos.size), C.c_uint32(position_stride))
    if rc == -2:
        raise RangeError(_err())
    if rc != 0:
        raise OracleError(_err())
    return out


def refit_tlas_instances(nodes, instance_indices, aabb_min, aabb_max, changed):
    """refitTlasInstances (packages/shared-bvh/src/tlas.ts:1141-1209);
    returns (refitted copy, affected node indices children-first)."""
    out = np.array(nodes, dtype=np.uint32).ravel().copy()
    perm = np.ascontiguousarray(instance_indices, dtype=np.uint32)
    mn = np.ascontiguousarray(aabb_min, dtype=np.float64).reshape(-1, 3)
    mx = np.ascontiguousarray(aabb_max, dtype=np.float64).reshape(-1, 3)
    ch = np.ascontiguousarray(changed, dtype=np.uint32)
    aff = np.zeros(max(out.size // 8, 1), np.uint32)
    n = C.c_uint32(0)
    rc = lib().orc_refit_tlas_instances(_p(out, U32P), C.c_uint32(out.size // 8), _p(perm, U32P),
                                        C.c_uint32(perm.size), C.c_uint32(mn.shape[0]), _p(mn, F64P),
                                        _p(mx, F64P), _p(ch, U32P), C.c_uint32(ch.size), _p(aff, U32P),
                                        C.byref(n))
    if rc != 0:
        raise OracleError(_err())
    return out, aff[: n.value].copy()


def js_min(a, b):
    """Math.min of two non-NaN numbers (a negative zero wins a tie)."""
    if a < b or b < a:
        return a if a < b else b
    return -0.0 if (a == 0 and (math.copysign(1, a) < 0 or math.copysign(1, b) < 0)) else a


def js_max(a, b):
    if a > b or b > a:
        return a if a > b else b
    return 0.0 if (a == 0 and not (math.copysign(1, a) < 0 and math.copysign(1, b) < 0)) else a


def local_aabb(positions4):
    """computeLocalAabb (scenePack.ts:172-195) of stride-4 positions, binary64."""
    mn, mx = [math.inf] * 3, [-math.inf] * 3
    for v in np.asarray(positions4, np.float32).reshape(-1, 4)[:, :3].astype(np.float64).tolist():
        for k in range(3):
            mn[k], mx[k] = js_min(mn[k], v[k]), js_max(mx[k], v[k])
    return mn, mx


def transform_aabb(l2w, mn, mx):
    """transformAabb (scenePack.ts:197-223): the 8 corners through the binary32
    column-major matrix in binary64 (transformPoint, :161-170)."""
    m = [float(x) for x in np.asarray(l2w, np.float32).ravel()]
    wmn, wmx = [math.inf] * 3, [-math.inf] * 3
    for c in range(8):
        p = [mx[0] if c & 1 else mn[0], mx[1] if c & 2 else mn[1], mx[2] if c & 4 else mn[2]]
        t = [m[r] * p[0] + m[4 + r] * p[1] + m[8 + r] * p[2] + m[12 + r] for r in range(4)]
        q = [t[0] / t[3], t[1] / t[3], t[2] / t[3]] if abs(t[3]) > 1e-8 else t[:3]
        for k in range(3):
            wmn[k], wmx[k] = js_min(wmn[k], q[k]), js_max(wmx[k], q[k])
    return wmn, wmx


def refit_primitive(pack, primitive, positions):
    """A packed scene after one mesh primitive's vertices move with its
    topology kept, as the reference's refit path does it
    (walkaround-hybrid HybridEnginePrimitiveUpdates.ts:1482-1560): the new
    positions replace the primitive's vertex slice, refitBvhBounds refits its
    BLAS, and refitTlasTransforms with localBoundsChanged (scenePack.ts:
    1310-1420) refits the TLAS over the primitive's instances from their new
    world boxes.  Returns (new pack, world_min, world_max) with the world
    bounds the binary64 union of every instance's box."""
    out = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in pack.items()}
    b = out["bindings"].reshape(-1, 8)
    rows = np.nonzero(b[:, 6] == primitive)[0]
    if rows.size != 1:
        raise OracleError(f"primitive {primitive} has no packed geometry")
    r = int(rows[0])
    root, _, vstart, vcount = (int(x) for x in b[r, :4])
    new = np.asarray(positions, np.float32).reshape(-1, 3)
    if new.shape[0] != vcount:
        raise OracleError("vertex count differs")
    pos = out["positions"].reshape(-1, 4)
    pos[vstart:vstart + vcount, :3] = new
    roots = sorted(int(x) for x in b[:, 0])
    total = out["bvh_nodes"].size // 8
    end = next((x for x in roots if x > root), total)
    nodes = out["bvh_nodes"].reshape(-1, 8)
    nodes[root:end] = refit_bvh_bounds(nodes[root:end], out["indices"], out["positions"]).reshape(-1, 8)
    lmn, lmx = local_aabb(pos[vstart:vstart + vcount])
    out["binding_aabb"].reshape(-1, 6)[r] = np.array(lmn + lmx, np.float32)
    inst_prim = out["instance_primitive"]
    n_inst = inst_prim.size
    boxes = {}
    for row in range(b.shape[0]):
        p, vs, vc = int(b[row, 6]), int(b[row, 2]), int(b[row, 3])
        boxes[p] = local_aabb(pos[vs:vs + vc])
    amn = np.zeros((n_inst, 3)); amx = np.zeros((n_inst, 3))
    l2w = out["tlas_l2w"].reshape(-1, 16)
    for i in range(n_inst):
        wmn, wmx = transform_aabb(l2w[i], *boxes[int(inst_prim[i])])
        amn[i], amx[i] = wmn, wmx
    changed = np.nonzero(inst_prim == primitive)[0]
    out["tlas_nodes"], _ = refit_tlas_instances(out["tlas_nodes"], out["tlas_instance_indices"], amn, amx, changed)
    return out, amn.min(axis=0), amx.max(axis=0)


def pack_scene(primitives):
    """packSceneFromCore (packages/shared-bvh/src/scenePack.ts:1104).

    primitives: list of dicts with keys kind ('mesh'|'instanced-mesh'),
    positions, normals, indices (optional), transform (mesh, optional) or
    instances (instanced-mesh), material_id.
    """
    keep = []
    arr = (Primitive * max(1, len(primitives)))()
    for i, p in enumerate(primitives):
        pos = np.ascontiguousarray(p["positions"], dtype=np.float32).ravel()
        nrm = np.ascontiguousarray(p["normals"], dtype=np.float32).ravel()
        idx = p.get("indices")
        idx = None if idx is None else np.ascontiguousarray(idx, dtype=np.uint32).ravel()
        if p["kind"] == "instanced-mesh":
            xf = np.ascontiguousarray(np.asarray(p["instances"], np.float32).reshape(-1, 16))
            kind, ntx = 1, xf.shape[0]
        else:
            t = p.get("transform")
            xf = None if t is None else np.ascontiguousarray(np.asarray(t, np.float32).reshape(16))
            kind, ntx = 0, (0 if xf is None else 1)
        keep += [pos, nrm, idx, xf]
        arr[i] = Primitive(kind, p.get("material_id", 0), _p(pos, F32P), pos.size, _p(nrm, F32P),
                           nrm.size, _p(idx, U32P), 0 if idx is None else idx.size,
                           _p(xf, F32P), ntx)
    handle = C.c_void_p()
    rc = lib().orc_pack_scene(arr, len(primitives), C.byref(handle))
    if rc != 0:
        raise OracleError(_err())
    out = {}
    try:
        for fid, name in enumerate(FIELDS):
            n = lib().orc_pack_field_size(handle, fid)
            buf = np.zeros(max(n, 1), np.float32 if name in FLOAT_FIELDS else np.uint32)
            lib().orc_pack_copy_field(handle, fid, buf.ctypes.data, C.c_uint64(buf.size))
            out[name] = buf[:n]
        out["warnings"] = lib().orc_pack_warning_count(handle)
    finally:
        lib().orc_pack_free(handle)
    return out


class PackedScene:
    """Holds numpy views of a packed scene and the orc_scene struct over them."""

    def __init__(self, pack):
        self.pack = pack
        self._keep = {k: np.ascontiguousarray(pack[k]) for k in (
            "positions", "indices", "bvh_nodes", "tlas_nodes", "tlas_instance_indices",
            "tlas_blas_roots", "tlas_w2l", "tlas_l2w")}
        k = self._keep
        self.struct = Scene(
            _p(k["positions"], F32P), k["positions"].size // 4,
            _p(k["indices"], U32P), k["indices"].size // 4,
            _p(k["bvh_nodes"], U32P), k["bvh_nodes"].size // 8,
            _p(k["tlas_nodes"], U32P), k["tlas_nodes"].size // 8,
            _p(k["tlas_instance_indices"], U32P), k["tlas_instance_indices"].size,
            _p(k["tlas_blas_roots"], U32P), _p(k["tlas_w2l"], F32P), _p(k["tlas_l2w"], F32P),
            k["tlas_blas_roots"].size)


def trace_closest(scene: PackedScene, rays, *, threads=None, counts=False):
    rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
    n = rays.size
    hits = np.zeros(n, HIT_DTYPE)
    cnt = np.zeros(n * 4, np.uint32) if counts else None
    threads = threads or min(16, os.cpu_count() or 1)
    lib().orc_trace_closest(C.byref(scene.struct), rays.ctypes.data_as(C.c_void_p), C.c_uint32(n),
                            hits.ctypes.data_as(C.c_void_p), _p(cnt, U32P), C.c_uint32(threads))
    return (hits, cnt.reshape(n, 4)) if counts else hits


def trace_any(scene: PackedScene, rays, *, threads=None):
    rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
    n = rays.size
    out = np.zeros(n, np.uint32)
    threads = threads or min(16, os.cpu_count() or 1)
    lib().orc_trace_any(C.byref(scene.struct), rays.ctypes.data_as(C.c_void_p), C.c_uint32(n),
                        _p(out, U32P), C.c_uint32(threads))
    return out


def single_sided_flags(scene: PackedScene, materials):
    """One byte per packed triangle: 1 when its material is opaque and not
    double-sided (materialAcceptsSidedHit, material.wgsl.ts:3259-3263);
    `materials` is the per-primitive material list of the scene (dicts with
    an optional doubleSided, default false)."""
    tri_mat = np.asarray(scene.pack["tri_material"], np.uint32)
    ds = np.array([1 if m.get("doubleSided", False) else 0 for m in materials] or [0], np.uint8)
    ok = tri_mat < ds.size
    out = np.ones(tri_mat.size, np.uint8)
    out[ok] = 1 - ds[tri_mat[ok]]
    return out


def trace_closest_sided(scene: PackedScene, single_sided, rays, *, threads=None, with_pos=False):
    """traceClosest (pt-webgpu intersection.wgsl.ts:462-490): back-face hits
    on single-sided materials are skipped by re-tracing from the hit."""
    rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
    n = rays.size
    hits = np.zeros(n, HIT_DTYPE)
    pos = np.zeros(3 * n, np.float32)
    flags = np.ascontiguousarray(single_sided, np.uint8)
    threads = threads or min(16, os.cpu_count() or 1)
    lib().orc_trace_closest_sided(C.byref(scene.struct), flags.ctypes.data_as(C.c_void_p),
                                  rays.ctypes.data_as(C.c_void_p), C.c_uint32(n),
                                  hits.ctypes.data_as(C.c_void_p), pos.ctypes.data_as(C.c_void_p),
                                  C.c_uint32(threads))
    return (hits, pos.reshape(n, 3)) if with_pos else hits


def trace_occluded_sided(scene: PackedScene, single_sided, rays, *, hit_limit=0xFFFFFFFF, threads=None):
    """traceSurfaceVisibility (connect.wgsl.ts:37-103), opaque materials:
    1 where an accepted surface blocks the ray."""
    rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
    n = rays.size
    out = np.zeros(n, np.uint32)
    flags = np.ascontiguousarray(single_sided, np.uint8)
    threads = threads or min(16, os.cpu_count() or 1)
    lib().orc_trace_occluded_sided(C.byref(scene.struct), flags.ctypes.data_as(C.c_void_p),
                                   C.c_uint32(hit_limit), rays.ctypes.data_as(C.c_void_p), C.c_uint32(n),
                                   _p(out, U32P), C.c_uint32(threads))
    return out


def orientation_sign(l2w):
    m = np.ascontiguousarray(l2w, np.float32).ravel()
    return int(lib().orc_orientation_sign(m.ctypes.data_as(F32P)))


HIT64_DTYPE = np.dtype([("hit", "<u4"), ("prim", "<u4"), ("inst", "<u4"), ("pad", "<u4"),
                        ("t", "<f8"), ("bary", "<f8", 3)])


class Cwbvh:
    """8-wide compressed BVH forest (buildCompressedWideBvhFromArrayBvh,
    packages/shared-bvh/src/compressedWideBvh.ts:248) over the binary trees at
    `roots` of a node array; arrays are copied out as numpy views."""

    def __init__(self, node_words, roots=(0,), tri_count=None):
        L = lib()
        L.orc_cwbvh_node_count.restype = C.c_uint32
        L.orc_cwbvh_node_count.argtypes = [C.c_void_p]
        L.orc_cwbvh_free.argtypes = [C.c_void_p]
        L.orc_cwbvh_stack_entries.restype = C.c_double
        words = np.ascontiguousarray(node_words, np.uint32).ravel()
        roots = np.ascontiguousarray(roots, np.uint32)
        if tri_count is None:
            tri_count = 1
        h = C.c_void_p()
        rc = L.orc_cwbvh_build(_p(words, U32P), C.c_uint32(words.size // 8), _p(roots, U32P),
                               C.c_uint32(roots.size), C.c_uint32(tri_count), C.byref(h))
        if rc != 0:
            raise OracleError(_err())
        self._h = h
        n = L.orc_cwbvh_node_count(h)
        self.node_count = n
        self.node_bounds = np.zeros((n, 6), np.float32)
        self.child_bounds = np.zeros((n, 8, 6), np.uint16)
        self.child_meta = np.zeros((n, 8, 3), np.uint32)
        self.child_count = np.zeros(n, np.uint32)
        self.roots = np.zeros(roots.size, np.uint32)
        self.status = np.zeros(roots.size, np.uint32)
        self.stack_entries = np.zeros(roots.size, np.float64)
        L.orc_cwbvh_copy(h, self.node_bounds.ctypes.data_as(F32P), self.child_bounds.ctypes.data_as(C.c_void_p),
                         self.child_meta.ctypes.data_as(U32P), self.child_count.ctypes.data_as(U32P),
                         self.roots.ctypes.data_as(U32P), self.status.ctypes.data_as(U32P),
                         self.stack_entries.ctypes.data_as(F64P))

    def close(self):
        if self._h:
            lib().orc_cwbvh_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def child_box(self, node, slot):
        """cwbvhChildBounds (compressedWideBvh.ts:481): binary64 decode."""
        pmn, pmx = self.node_bounds[node, :3].astype(np.float64), self.node_bounds[node, 3:].astype(np.float64)
        q = self.child_bounds[node, slot].astype(np.float64)

        def deq(qv, mn, mx):
            if mx == mn or qv <= 0:
                return mn
            if qv >= 65535:
                return mx
            t = qv / 65535
            return mn * (1 - t) + mx * t

        return ([deq(q[a], pmn[a], pmx[a]) for a in range(3)], [deq(q[3 + a], pmn[a], pmx[a]) for a in range(3)])


def cwbvh_stack_entries(child_meta, child_count, root=0):
    """requiredCwbvhTraversalStackEntries (compressedWideBvh.ts:60)."""
    L = lib()
    L.orc_cwbvh_stack_entries.restype = C.c_double
    m = np.ascontiguousarray(child_meta, np.uint32).ravel()
    c = np.ascontiguousarray(child_count, np.uint32).ravel()
    return L.orc_cwbvh_stack_entries(_p(m, U32P), _p(c, U32P), C.c_uint32(c.size), C.c_uint32(root))


def scene_cwbvh(scene: PackedScene):
    """The wide forest of a packed scene: one tree per distinct BLAS root;
    returns (Cwbvh, wide root per packed instance)."""
    blas_roots = np.asarray(scene.pack["tlas_blas_roots"], np.uint32)
    uniq = np.unique(blas_roots)
    cw = Cwbvh(scene.pack["bvh_nodes"], uniq, tri_count=scene.pack["indices"].size // 4)
    lut = dict(zip(uniq.tolist(), cw.roots.tolist()))
    return cw, np.array([lut[int(r)] for r in blas_roots], np.uint32)


def cwbvh_trace_closest(scene: PackedScene, cw, inst_roots, rays, *, threads=None):
    """intersectCompressedWideBvhFirstHit (compressedWideBvh.ts:754) per
    tlasIntersect candidate (tlas.ts:1223), binary64."""
    rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
    out = np.zeros(rays.size, HIT64_DTYPE)
    roots = np.ascontiguousarray(inst_roots, np.uint32)
    threads = threads or min(16, os.cpu_count() or 1)
    rc = lib().orc_cwbvh_trace_closest(C.byref(scene.struct), cw._h, _p(roots, U32P),
                                       rays.ctypes.data_as(C.c_void_p), C.c_uint32(rays.size),
                                       out.ctypes.data_as(C.c_void_p), C.c_uint32(threads))
    if rc != 0:
        raise OracleError(_err())
    return out


class CwbvhView(C.Structure):
    _fields_ = [("node_bounds", F32P), ("child_bounds", C.c_void_p), ("child_meta", U32P),
                ("child_count", U32P), ("node_count", C.c_uint32), ("inst_root", U32P)]


def trace_closest_cwbvh(scene: PackedScene, cw, inst_roots, rays, *, threads=None, single_sided=None):
    """pt-webgpu's `cwbvh-closest` mode in binary32 (intersection.wgsl.ts:
    576-838): Moller-Trumbore proposal in the wide forest, watertight
    re-verification, binary fallback.  Same hit record as trace_closest."""
    rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
    hits = np.zeros(rays.size, HIT_DTYPE)
    roots = np.ascontiguousarray(inst_roots, np.uint32)
    keep = [np.ascontiguousarray(a) for a in (cw.node_bounds, cw.child_bounds, cw.child_meta, cw.child_count)]
    v = CwbvhView(keep[0].ctypes.data_as(F32P), keep[1].ctypes.data_as(C.c_void_p), keep[2].ctypes.data_as(U32P),
                  keep[3].ctypes.data_as(U32P), cw.node_count, roots.ctypes.data_as(U32P))
    threads = threads or min(16, os.cpu_count() or 1)
    flags = None if single_sided is None else np.ascontiguousarray(single_sided, np.uint8)
    lib().orc_trace_closest_cwbvh_sided(C.byref(scene.struct), C.byref(v),
                                        None if flags is None else flags.ctypes.data_as(C.c_void_p),
                                        rays.ctypes.data_as(C.c_void_p), C.c_uint32(rays.size),
                                        hits.ctypes.data_as(C.c_void_p), C.c_uint32(threads))
    return hits


LT_FLOATS = 16
LT_BUCKETS = 1 << 24


def light_tree_build(powers, centroids, aabb_min, aabb_max, cones=None):
    """buildLightTree + packLightTreeForGPU (packages/shared-samplers/src/
    lightTree.ts:492, :611).  cones: per emitter None (full sphere) or
    (axis, theta_o, theta_e).  Returns the packed (nodes, 16) float32 array."""
    n = len(powers)
    pw = np.ascontiguousarray(powers, np.float64)
    c = np.ascontiguousarray(centroids, np.float64).reshape(n, 3)
    mn = np.ascontiguousarray(aabb_min, np.float64).reshape(n, 3)
    mx = np.ascontiguousarray(aabb_max, np.float64).reshape(n, 3)
    has = np.zeros(n, np.uint8)
    axis = np.zeros((n, 3), np.float64)
    to = np.zeros(n, np.float64)
    te = np.zeros(n, np.float64)
    for i, cn in enumerate(cones or [None] * n):
        if cn is not None:
            has[i] = 1
            axis[i] = cn[0]
            to[i] = cn[1] if len(cn) > 1 and cn[1] is not None else 0.0
            te[i] = cn[2] if len(cn) > 2 and cn[2] is not None else np.pi / 2
    out = np.zeros((max(1, 2 * n - 1), LT_FLOATS), np.float32)
    count = C.c_uint32(0)
    rc = lib().orc_light_tree_build(C.c_uint32(n), _p(pw, F64P), _p(c, F64P), _p(mn, F64P), _p(mx, F64P),
                                    has.ctypes.data_as(C.c_void_p), _p(axis, F64P), _p(to, F64P), _p(te, F64P),
                                    _p(out, F32P), C.byref(count))
    if rc != 0:
        raise OracleError(_err())
    return out[:count.value].copy()


def light_tree_pack_nodes(nodes):
    """packLightTreeForGPU over explicit nodes: dicts with emitter, power,
    left, right, aabb_min, aabb_max and optional axis/theta_o/theta_e."""
    n = len(nodes)
    em = np.array([d["emitter"] for d in nodes], np.int32)
    pw = np.array([d["power"] for d in nodes], np.float64)
    lf = np.array([d["left"] for d in nodes], np.int32)
    rt = np.array([d["right"] for d in nodes], np.int32)
    mn = np.array([d.get("aabb_min", (0, 0, 0)) for d in nodes], np.float64)
    mx = np.array([d.get("aabb_max", (0, 0, 0)) for d in nodes], np.float64)
    ax = np.array([d.get("axis", (0, 0, 0)) for d in nodes], np.float64)
    to = np.array([d.get("theta_o", np.pi) for d in nodes], np.float64)
    te = np.array([d.get("theta_e", np.pi) for d in nodes], np.float64)
    out = np.zeros((n, LT_FLOATS), np.float32)
    rc = lib().orc_light_tree_pack_nodes(C.c_uint32(n), em.ctypes.data_as(C.c_void_p), _p(pw, F64P),
                                         lf.ctypes.data_as(C.c_void_p), rt.ctypes.data_as(C.c_void_p),
                                         _p(mn, F64P), _p(mx, F64P), _p(ax, F64P), _p(to, F64P), _p(te, F64P),
                                         _p(out, F32P))
    if rc != 0:
        raise OracleError(_err())
    return out


def _lt_setup():
    L = lib()
    L.orc_light_tree_importance.restype = C.c_float
    L.orc_light_tree_importance.argtypes = [F32P, C.c_uint32, F32P, C.c_float]
    L.orc_light_tree_sample.restype = C.c_int32
    L.orc_light_tree_sample.argtypes = [F32P, C.c_uint32, F32P, C.c_float, C.c_uint32, F32P]
    L.orc_light_tree_pdf.restype = C.c_float
    L.orc_light_tree_pdf.argtypes = [F32P, C.c_uint32, F32P, C.c_float, C.c_int32]
    return L


def light_tree_importance(packed, node, p, floor_d2):
    pk = np.ascontiguousarray(packed, np.float32)
    pt = np.ascontiguousarray(p, np.float32)
    return _lt_setup().orc_light_tree_importance(_p(pk, F32P), node, _p(pt, F32P), floor_d2)


def light_tree_sample(packed, p, floor_d2, bucket):
    """samplePackedLightTreeCPU with the 24-bit root draw `bucket`."""
    pk = np.ascontiguousarray(packed, np.float32)
    pt = np.ascontiguousarray(p, np.float32)
    pdf = np.zeros(1, np.float32)
    e = _lt_setup().orc_light_tree_sample(_p(pk, F32P), pk.shape[0], _p(pt, F32P), floor_d2, int(bucket),
                                          _p(pdf, F32P))
    return int(e), float(pdf[0])


def light_tree_pdf(packed, p, floor_d2, emitter):
    pk = np.ascontiguousarray(packed, np.float32)
    pt = np.ascontiguousarray(p, np.float32)
    return float(_lt_setup().orc_light_tree_pdf(_p(pk, F32P), pk.shape[0], _p(pt, F32P), floor_d2, emitter))


def fp64_hit_params(scene: PackedScene, rays, hits):
    rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
    hits = np.ascontiguousarray(hits, dtype=HIT_DTYPE)
    out = np.zeros((rays.size, 3), np.float64)
    lib().orc_fp64_hit_params(C.byref(scene.struct), rays.ctypes.data_as(C.c_void_p),
                              hits.ctypes.data_as(C.c_void_p), C.c_uint32(rays.size),
                              out.ctypes.data_as(F64P))
    return out


def watertight(o, d, a, b, c, exclusive_min_t=0.0):
    args = [np.ascontiguousarray(v, dtype=np.float32) for v in (o, d, a, b, c)]
    out = np.zeros(10, np.float32)
    lib().orc_watertight(*[_p(v, F32P) for v in args], C.c_float(exclusive_min_t), _p(out, F32P))
    return {"hit": bool(out[0]), "t": out[1], "point": out[2:5].copy(), "bary": out[5:8].copy(),
            "side": out[8], "zero_edge_mask": int(out[9])}


def pcg_init(px, py, seed):
    return lib().orc_pcg_init(px, py, seed)


def pcg_stream(px, py, seed, n):
    s = C.c_uint32(pcg_init(px, py, seed))
    return [lib().orc_pcg_next(C.byref(s)) for _ in range(n)]


# ---------------------------------------------------------------------------
# Integrator oracle (integrator.c)
# ---------------------------------------------------------------------------

class RenderDesc(C.Structure):
    _fields_ = [
        ("scene", C.POINTER(Scene)), ("normals", F32P), ("tri_material", U32P),
        ("materials", F32P), ("material_count", C.c_uint32), ("emitters", F32P),
        ("emitter_count", C.c_uint32), ("env_present", C.c_uint32),
        ("env_w", C.c_uint32), ("env_h", C.c_uint32),
        ("env_cdf", F32P), ("env_texels", F32P), ("env_rot", C.c_float * 2),
        ("inv_view_proj", C.c_float * 16),
        ("width", C.c_uint32), ("height", C.c_uint32), ("frame_seed", C.c_uint32),
        ("max_bounces", C.c_uint32), ("ray_bias", C.c_float), ("sampling", C.c_uint32),
        ("light_tree", F32P), ("light_tree_nodes", C.c_uint32), ("aux", F64P),
        ("lights", F32P), ("n_dir", C.c_uint32), ("n_point", C.c_uint32), ("n_spot", C.c_uint32),
        ("n_rect", C.c_uint32),
    ]


def material_table(materials):
    """24 floats per material: base rgb, roughness, metallic, Le rgb
    (emissive x emissiveIntensity), doubleSided, then the extension lobes as
    materialPacking.ts:259-316 packs them (clamped to [0, 1]): clearcoat,
    clearcoatRoughness, sheen, sheenRoughness, sheenColor rgb, anisotropy, the
    cosine and sine of anisotropyRotation (binary64, rounded once), then
    specularColor rgb (>= 0), specularIntensity, 1 pad."""
    out = np.zeros((max(1, len(materials)), 24), np.float32)
    c01 = lambda x: min(1.0, max(0.0, float(np.float32(x))))
    for i, m in enumerate(materials):
        # the engine receives f32 emissive and intensity; Le is their binary64
        # product rounded once to f32
        em = np.asarray(m.get("emissive", [0, 0, 0]), np.float32).astype(np.float64)
        inten = float(np.float32(m.get("emissiveIntensity", 1.0 if em.any() else 0.0)))
        out[i, 0:3] = m.get("baseColor", [0.8, 0.8, 0.8])
        out[i, 3] = m.get("roughness", 1.0)
        out[i, 4] = m.get("metallic", 0.0)
        out[i, 5:8] = (em * inten).astype(np.float32)
        out[i, 8] = 1.0 if m.get("doubleSided", False) else 0.0
        out[i, 9] = c01(m.get("clearcoat", 0.0))
        out[i, 10] = c01(m.get("clearcoatRoughness", 0.0))
        out[i, 11] = c01(m.get("sheen", 0.0))
        out[i, 12] = c01(m.get("sheenRoughness", 0.0))
        out[i, 13:16] = [c01(c) for c in m.get("sheenColor", [0.0, 0.0, 0.0])]
        out[i, 16] = c01(m.get("anisotropy", 0.0))
        rot = float(np.float32(m.get("anisotropyRotation", 0.0)))
        out[i, 17], out[i, 18] = math.cos(rot), math.sin(rot)
        out[i, 19:22] = [max(0.0, float(np.float32(c))) for c in m.get("specularColor", [1.0, 1.0, 1.0])]
        out[i, 22] = c01(m.get("specularIntensity", 1.0))
    if not materials:
        out[0, 0:4] = [0.8, 0.8, 0.8, 1.0]
        out[0, 17] = 1.0
        out[0, 19:23] = 1.0
    return out


def _bsdf_lib():
    L = lib()
    if not getattr(L, "_bsdf_ready", False):
        D3P = C.POINTER(C.c_double)
        L.orc_bsdf_eval.argtypes = [F32P, D3P, D3P, D3P, D3P]
        L.orc_bsdf_eval.restype = None
        L.orc_bsdf_pdf.argtypes = [F32P, D3P, D3P, D3P]
        L.orc_bsdf_pdf.restype = C.c_double
        L.orc_bsdf_sample.argtypes = [F32P, F32P, F32P, U32P, F32P, U32P]
        L.orc_bsdf_sample.restype = None
        L._bsdf_ready = True
    return L


def bsdf_eval(mat_row, n, wo, wi):
    """evaluateBrdfFullWithClearcoatNormal (the kernel without N.L) of one
    material_table row."""
    q = np.ascontiguousarray(mat_row, np.float32)
    v = [np.ascontiguousarray(x, np.float64) for x in (n, wo, wi)]
    out = np.zeros(3)
    _bsdf_lib().orc_bsdf_eval(_p(q, F32P), *[_p(x, F64P) for x in v], _p(out, F64P))
    return out


def bsdf_pdf(mat_row, n, wo, wi):
    """brdfDirectionalPdfFullWithClearcoatNormal of one material_table row."""
    q = np.ascontiguousarray(mat_row, np.float32)
    v = [np.ascontiguousarray(x, np.float64) for x in (n, wo, wi)]
    return _bsdf_lib().orc_bsdf_pdf(_p(q, F32P), *[_p(x, F64P) for x in v])


def bsdf_sample(mat_row, n, wo, count, seed=1):
    """`count` sampled directions (f32) and lobes (0 diffuse, 1 specular,
    2 clearcoat, 3 sheen) from one PCG stream."""
    q = np.ascontiguousarray(mat_row, np.float32)
    nf = np.ascontiguousarray(n, np.float32)
    wf = np.ascontiguousarray(wo, np.float32)
    st = np.array([seed], np.uint32)
    wi = np.zeros((count, 3), np.float32)
    lobe = np.zeros(count, np.uint32)
    L = _bsdf_lib()
    w = np.zeros(3, np.float32)
    lb = np.zeros(1, np.uint32)
    for k in range(count):
        L.orc_bsdf_sample(_p(q, F32P), _p(nf, F32P), _p(wf, F32P), _p(st, U32P), _p(w, F32P), _p(lb, U32P))
        wi[k] = w
        lobe[k] = lb[0]
    return wi, lobe


def build_emitters(pack, mat_table):
    """Mesh-area emitter records, one per emissive triangle
    (pt-webgpu/src/scene/emitterPacking.ts:748-820): primitives in scene
    order, each primitive's packed instances in order, source triangles in
    order.  World vertices are the binary64 local-to-world image of the f32
    local vertices, rounded to f32.  16 floats per record: a, b, c, Le,
    twoSided, packed instance, packed triangle, 0."""
    b = pack["bindings"].reshape(-1, 8)
    pos = pack["positions"].reshape(-1, 4)[:, :3].astype(np.float64)
    idx = pack["indices"].reshape(-1, 4)[:, :3]
    tri_src = pack["tri_source"]
    inst_prim = pack["instance_primitive"]
    l2w = pack["tlas_l2w"].reshape(-1, 16).astype(np.float64).tolist()
    tri_mat = pack["tri_material"]
    recs = []
    for blas_root, n_inst, vs, vc, ts, tc, prim, kind in b.tolist():
        mat = int(tri_mat[ts])
        le = mat_table[mat, 5:8]
        if not (le > 0).any():
            continue
        two = mat_table[mat, 8]
        # packed triangle of each source triangle
        order = np.empty(tc, np.int64)
        order[tri_src[ts:ts + tc]] = np.arange(ts, ts + tc)
        for inst in np.nonzero(inst_prim == prim)[0]:
            m = l2w[inst]
            for s in range(tc):
                t = int(order[s])
                w = []
                for k in range(3):
                    # transformPoint (pt-webgpu/src/math/mat4.ts:20-33), left
                    # to right in binary64, divided only when |w| > 1e-8
                    x, y, z = (float(v) for v in pos[idx[t, k]])
                    q = [m[j] * x + m[4 + j] * y + m[8 + j] * z + m[12 + j] for j in range(4)]
                    if abs(q[3]) > 1e-8:
                        q = [q[0] / q[3], q[1] / q[3], q[2] / q[3]]
                    w.append(np.array(q[:3], np.float64).astype(np.float32))
                recs.append(np.concatenate([w[0], w[1], w[2], le, [two, inst, t, 0]]).astype(np.float32))
    if not recs:
        return np.zeros((0, 16), np.float32)
    return np.stack(recs)


def _hypot3_js(a, b, c):
    """Math.hypot (V8): max-scaled, compensated sum of squares."""
    v = [abs(float(a)), abs(float(b)), abs(float(c))]
    m = max(v)
    if m == 0 or np.isinf(m):
        return m
    s = comp = 0.0
    for x in v:
        n = x / m
        t = n * n - comp
        nxt = s + t
        comp = (nxt - s) - t
        s = nxt
    return np.sqrt(s) * m


def _area_vector_f32(u, v, coefficient):
    """classifyAreaVectorF32 (pt-webgpu/src/scene/areaEmitterGeometry.ts:
    51-128): coefficient * |u x v| in f32 arithmetic; 0 when degenerate."""
    f = np.float32
    u = [f(x) for x in u]
    v = [f(x) for x in v]
    es = max(abs(x) for x in u + v)
    if not (es > 0 and np.isfinite(es)):
        return f(0)
    su = [x / es for x in u]
    sv = [x / es for x in v]
    cx = su[1] * sv[2] - su[2] * sv[1]
    cy = su[2] * sv[0] - su[0] * sv[2]
    cz = su[0] * sv[1] - su[1] * sv[0]
    cs = max(abs(cx), abs(cy), abs(cz))
    if not (cs > 0 and np.isfinite(cs)):
        return f(0)
    dl = f(_hypot3_js(cx / cs, cy / cs, cz / cs))
    if not (dl > 0):
        return f(0)
    with np.errstate(over="ignore"):
        area = ((f(coefficient) * (cs * dl)) * es) * es
        inv = f(1) / area if area > 0 else f(0)
    return area if (area > 0 and np.isfinite(area) and inv > 0 and np.isfinite(inv)) else f(0)


def _lum_power(rad, area=None):
    """emitterPower (pt-webgpu/src/scene/emitterPacking.ts:900-934): f32
    luminance of the f32 radiance, times the f32 area of a finite emitter."""
    lum = np.float32(max(0.0, 0.2126 * float(rad[0]) + 0.7152 * float(rad[1]) + 0.0722 * float(rad[2])))
    if area is None:
        return float(lum)
    return float(np.float32(float(lum) * float(area)))


def _normalize_f32(v):
    """normalizeDirectionF32 (areaEmitterGeometry.ts:181-198), or None."""
    s = max(abs(float(x)) for x in v)
    if not (s > 0 and math.isfinite(s)):
        return None
    q = [float(x) / s for x in v]
    ln = _hypot3_js(*q)
    if not (ln > 0):
        return None
    return [float(np.float32(x / ln)) for x in q]


def pack_analytic(emitters):
    """packEmitterArrays (pt-webgpu/src/scene/emitterPacking.ts:600-760) for
    the analytic kinds, as records of 16 f32 in light-selection order --
    directional (toward-light, angular diameter, irradiance, no-shadow),
    point (position, no-shadow, radiance, 0, cutoff, decay), spot (position,
    no-shadow, axis, cos outer, radiance, cos inner, cutoff, decay), then
    rect and disc (centre, no-shadow, u, 0, v, 0, radiance, shape) -- and the
    counts (n_dir, n_point, n_spot, n_rect).  Raises ValueError where the
    reference throws RangeError."""
    f = np.float32
    recs = {k: [] for k in ("directional", "point", "spot", "rect-area", "disc-area")}
    for e in emitters:
        recs[e["kind"]].append(e)

    def rad(e):
        r = [float(f(f(c) * f(e.get("intensity", 1.0)))) for c in e["color"]]
        if not all(math.isfinite(x) and x >= 0 for x in r):
            raise ValueError("radiance must be finite and non-negative")
        return r

    def shadow(e):
        return 1.0 if e.get("castShadow", True) is False else 0.0

    out = []
    for e in recs["directional"]:
        x, y, z = (float(c) for c in e["direction"])
        ln = _hypot3_js(x, y, z)
        t = [0.0, 1.0, 0.0] if ln < 1e-8 else [-x / ln, -y / ln, -z / ln]
        ad = e.get("angularDiameter", 0.0)
        ad = ad if (ad is not None and math.isfinite(ad) and ad > 0) else 0.0
        out.append(t + [ad] + rad(e) + [shadow(e)] + [0.0] * 8)
    for e in recs["point"]:
        dist = e.get("distance", 0.0)
        dist = dist if dist and dist > 0 else 0.0
        out.append(list(e["position"]) + [shadow(e)] + rad(e) + [0.0, dist, e.get("decay", 2.0)] + [0.0] * 6)
    for e in recs["spot"]:
        x, y, z = (float(c) for c in e["direction"])
        ln = _hypot3_js(x, y, z)
        d = [0.0, -1.0, 0.0] if ln < 1e-8 else [x / ln, y / ln, z / ln]
        pen = min(1.0, max(0.0, e.get("penumbra", 0.0) or 0.0))
        outer = e["angle"]
        inner = outer * (1 - pen)
        dist = e.get("distance", 0.0)
        dist = dist if dist and dist > 0 else 0.0
        out.append(list(e["position"]) + [shadow(e)] + d + [math.cos(outer)] + rad(e) + [math.cos(inner)] +
                   [dist, e.get("decay", 2.0), 0.0, 0.0])
    for e in recs["rect-area"]:
        u = [float(f(c)) for c in e["uAxis"]]
        v = [float(f(c)) for c in e["vAxis"]]
        if not _area_vector_f32(u, v, 4.0) > 0:
            raise ValueError("rect-area emitter does not retain positive f32 area")
        out.append(list(e["position"]) + [shadow(e)] + u + [0.0] + v + [0.0] + rad(e) + [0.0])
    for e in recs["disc-area"]:
        r = e["radius"]
        if not (math.isfinite(r) and r > 0):
            raise ValueError("disc-area emitter has a non-positive radius")
        n = _normalize_f32(e["normal"])
        if n is None:
            raise ValueError("disc-area emitter has a degenerate normal")
        ax, ay, az = (1.0, 0.0, 0.0) if abs(n[1]) > 0.999 else (0.0, 1.0, 0.0)
        ux, uy, uz = n
        t = _normalize_f32([ay * uz - az * uy, az * ux - ax * uz, ax * uy - ay * ux])
        if t is None:
            raise ValueError("disc-area emitter has a degenerate tangent basis")
        b = [uy * t[2] - uz * t[1], uz * t[0] - ux * t[2], ux * t[1] - uy * t[0]]
        uax = [float(f(c * r)) for c in t]
        vax = [float(f(c * r)) for c in b]
        if not _area_vector_f32(uax, vax, math.pi) > 0:
            raise ValueError("disc-area emitter does not retain positive f32 area")
        out.append(list(e["position"]) + [shadow(e)] + uax + [0.0] + vax + [0.0] + rad(e) + [1.0])
    counts = (len(recs["directional"]), len(recs["point"]), len(recs["spot"]),
              len(recs["rect-area"]) + len(recs["disc-area"]))
    arr = np.asarray(out, np.float32).reshape(-1, 16)
    return arr, counts


def _analytic_leaves(records, counts):
    """Leaves of the positional analytic lights (flatEmitterWalk.ts:115-188 as
    emitterPacking.ts:1090-1140 uses it): point / spot at their position
    (spot: a cone of half angle acos(cos outer) about its axis), rect / disc
    on their box with a hemisphere about u x v."""
    nd, npt, ns, nr = counts
    out = []
    for k in range(nd, nd + npt + ns + nr):
        q = [float(x) for x in records[k]]
        p = q[0:3]
        if k < nd + npt + ns:
            spot = k >= nd + npt
            rad = q[8:11] if spot else q[4:7]
            cone = None
            if spot:
                ln = _hypot3_js(*q[4:7])
                axis = [x / ln for x in q[4:7]] if ln > 1e-12 else [0.0, 0.0, 0.0]
                cone = (axis, 0.0, math.acos(max(-1.0, min(1.0, q[7]))))
            out.append((_lum_power(rad), list(p), list(p), list(p), cone))
        else:
            u, v = q[4:7], q[8:11]
            disc = abs(q[15] - 1.0) < 0.5
            area = _area_vector_f32(u, v, math.pi if disc else 4.0)
            if disc:
                r = _hypot3_js(*u)
                ext = [r, r, r]
            else:
                ext = [abs(u[i]) + abs(v[i]) for i in range(3)]
            lo = [p[i] - ext[i] for i in range(3)]
            hi = [p[i] + ext[i] for i in range(3)]
            n = [u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]]
            ln = _hypot3_js(*n)
            axis = [x / ln for x in n] if ln > 1e-12 else [0.0, 0.0, 0.0]
            out.append((_lum_power(q[12:15], area), list(p), lo, hi, (axis, 0.0, math.pi / 2)))
    return out


def _triangle_area_f32(a, b, c):
    """measureTriangleAreaF32 (pt-webgpu/src/scene/areaEmitterGeometry.ts:
    37-173): equilibrated f32 cross product, one rounding per operation;
    0 when degenerate."""
    f = np.float32
    u = [f(b[k]) - f(a[k]) for k in range(3)]
    v = [f(c[k]) - f(a[k]) for k in range(3)]
    es = max(abs(x) for x in u + v)
    if not (es > 0 and np.isfinite(es)):
        return f(0)
    su = [x / es for x in u]
    sv = [x / es for x in v]
    cx = su[1] * sv[2] - su[2] * sv[1]
    cy = su[2] * sv[0] - su[0] * sv[2]
    cz = su[0] * sv[1] - su[1] * sv[0]
    cs = max(abs(cx), abs(cy), abs(cz))
    if not (cs > 0 and np.isfinite(cs)):
        return f(0)
    dl = f(_hypot3_js(cx / cs, cy / cs, cz / cs))
    if not (dl > 0):
        return f(0)
    area = ((f(0.5) * (cs * dl)) * es) * es
    with np.errstate(divide="ignore", over="ignore"):
        inv = f(1) / area
    return area if (area > 0 and np.isfinite(area) and inv > 0 and np.isfinite(inv)) else f(0)


def scene_light_tree(emitters, env_radiance=None, env_size=(64, 32), analytic=None, env_map=None):
    """The scene's NEE light tree (emitterPacking.ts:1005-1230 leaves, then
    light_tree_build): one leaf per mesh emitter record in order -- power =
    f32 luminance of its radiance x its f32 area, bounds of its vertices,
    an outward hemisphere cone unless two-sided -- then the environment
    (power = luminance x solid angle summed over the map, bounds = the union
    of the emitters' bounds).  None below two lights."""
    em = np.asarray(emitters, np.float32).reshape(-1, 16)
    env = environment_table(env_radiance, env_size, env_map)
    env_present = env is not None
    records, counts = analytic if analytic is not None else (np.zeros((0, 16), np.float32), (0, 0, 0, 0))
    n = em.shape[0] + (1 if env_present else 0) + sum(counts)
    if n < 2:
        return None
    powers, cents, lo, hi, cones = [], [], [], [], []
    for pw, c, l, h, cone in _analytic_leaves(records, counts):
        powers.append(pw)
        cents.append(c)
        lo.append(np.asarray(l, np.float64))
        hi.append(np.asarray(h, np.float64))
        cones.append(cone)
    for r in em:
        a, b, c = r[0:3], r[3:6], r[6:9]
        le = r[9:12]
        lum = np.float32(max(0.0, 0.2126 * float(le[0]) + 0.7152 * float(le[1]) + 0.0722 * float(le[2])))
        area = _triangle_area_f32(a, b, c)
        powers.append(float(np.float32(float(lum) * float(area))))
        A, B, Cc = (x.astype(np.float64) for x in (a, b, c))
        cents.append([(A[k] + B[k] + Cc[k]) / 3 for k in range(3)])
        lo.append(np.minimum(np.minimum(A, B), Cc))
        hi.append(np.maximum(np.maximum(A, B), Cc))
        if r[12] > 0.5:
            cones.append(None)
        else:
            e1, e2 = B - A, Cc - A
            nrm = [e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]]
            ln = _hypot3_js(*nrm)
            cones.append(([x / ln for x in nrm] if ln > 1e-12 else [0.0, 0.0, 0.0], 0.0, np.pi / 2))
    if env_present:
        total = env["power"]
        pw = float(np.float32(total))
        env_power = (pw if pw > 0 else 2.0 ** -149) if total > 0 else 0.0
    # directional lights and the environment sit on the union of the
    # positional lights' bounds (the origin when there are none)
    if lo:
        ulo, uhi = np.min(np.array(lo), axis=0), np.max(np.array(hi), axis=0)
    else:
        ulo = uhi = np.zeros(3)
    ucen = [(ulo[k] + uhi[k]) / 2 for k in range(3)]
    nd = counts[0]
    for k in range(nd - 1, -1, -1):
        powers.insert(0, _lum_power([float(x) for x in records[k][4:7]]))
        cents.insert(0, ucen)
        lo.insert(0, ulo)
        hi.insert(0, uhi)
        cones.insert(0, None)
    if env_present:
        powers.append(env_power)
        lo.append(ulo)
        hi.append(uhi)
        cents.append(ucen)
        cones.append(None)
    return light_tree_build(powers, cents, lo, hi, cones)


def represented_cdf(weights, bucket_bits=24, counts=False):
    """buildRepresentedDistributionF32 (shared-samplers/src/representedDistribution.ts:
    44-175): Hamilton apportionment of 2^bucket_bits buckets, f32 cdf (and the
    bucket counts when `counts`)."""
    w = np.asarray(weights, np.float64)
    total = 1 << bucket_bits
    pos_idx = [i for i in range(w.size) if w[i] > 0]
    buckets = np.zeros(w.size, np.int64)
    if pos_idx:
        mx = max(w[i] for i in pos_idx)
        s, comp = 0.0, 0.0
        for i in pos_idx:
            sc = w[i] / mx
            corr = sc - comp
            nxt = s + corr
            comp = (nxt - s) - corr
            s = nxt
        remaining = total - len(pos_idx)
        rema = []
        allocated = len(pos_idx)
        for i in pos_idx:
            quota = 0.0 if remaining == 0 else ((w[i] / mx) / s) * remaining
            extra = int(np.floor(quota))
            buckets[i] = 1 + extra
            allocated += extra
            rema.append((i, quota - extra))
        rema.sort(key=lambda r: (-r[1], r[0]))
        deficit = total - allocated
        cur = 0
        while deficit > 0:
            buckets[rema[cur % len(rema)][0]] += 1
            cur += 1
            deficit -= 1
        if deficit < 0:
            asc = sorted(rema, key=lambda r: (r[1], -r[0]))
            excess, cur = -deficit, 0
            while excess > 0:
                i = asc[cur % len(asc)][0]
                cur += 1
                if buckets[i] <= 1:
                    continue
                buckets[i] -= 1
                excess -= 1
    cdf = np.zeros(w.size + 1, np.float32)
    cum = np.cumsum(buckets)
    cdf[1:] = (cum / float(total)).astype(np.float32)
    return (cdf, buckets) if counts else cdf


def _luminance(r, g, b):
    return 0.2126 * r + 0.7152 * g + 0.0722 * b


def environment_table(env_radiance=None, env_size=(64, 32), env_map=None):
    """The environment's importance table and texels
    (pt-webgpu/src/scene/environmentPacking.ts:283-440), or None for no
    environment or a map that is black everywhere.  `env_map` = {"texels":
    (H, W, C) float32 linear radiance, C in 1..4 as environmentMapPixels.ts:
    138-142 expands it, "intensity", "rotation_y"}; otherwise `env_radiance`
    is a constant [r, g, b] over an env_size map.  Weights = luminance of the
    unscaled texel x its solid angle; represented 24-bit CDF; per texel the
    binary32 intensity-scaled rgb and pdf = effective pmf / solid angle;
    power = luminance x solid angle of the scaled map; rotation rounded to
    f32 after reduction mod 2 pi (environmentRotation.ts:10-26), cos and sin
    rounded once."""
    if env_map is not None:
        tex = np.asarray(env_map["texels"], np.float32)
        if tex.ndim == 2:
            tex = tex[:, :, None]
        H, W, ch = tex.shape
        rgb = np.stack([tex[:, :, 0], tex[:, :, 1] if ch > 1 else tex[:, :, 0],
                        tex[:, :, 2] if ch > 2 else tex[:, :, 0]], axis=-1).reshape(-1, 3)
        intensity = np.float32(env_map.get("intensity", 1.0))
        rot = float(np.float32(math.fmod(float(np.float32(env_map.get("rotation_y", 0.0))), 2 * math.pi)))
    else:
        if env_radiance is None or not any(v > 0 for v in env_radiance):
            return None
        W, H = env_size
        rgb = np.tile(np.asarray(env_radiance, np.float32)[:3], (W * H, 1))
        intensity, rot = np.float32(1.0), 0.0
    dphi = 2 * math.pi / W
    # libm cosines, as the host code's
    rows = np.array([dphi * (math.cos(j / H * math.pi) - math.cos((j + 1) / H * math.pi)) for j in range(H)])
    omega = np.repeat(rows, W)
    r64 = rgb.astype(np.float64)
    weights = np.maximum(0.0, _luminance(r64[:, 0], r64[:, 1], r64[:, 2]) * omega)
    total = float(np.sum(weights))
    if not total > 0:
        return None
    scaled = (rgb * intensity).astype(np.float32)  # one binary32 product per channel
    s64 = scaled.astype(np.float64)
    # summed in texel order, as the host code sums
    power = float(np.cumsum(np.maximum(0.0, _luminance(s64[:, 0], s64[:, 1], s64[:, 2]) * omega))[-1])
    cdf = represented_cdf(weights)
    pdf = ((cdf[1:].astype(np.float64) - cdf[:-1].astype(np.float64)) / omega).astype(np.float32)
    texels = np.concatenate([scaled, pdf[:, None]], axis=1).astype(np.float32)
    return {"w": W, "h": H, "cdf": cdf, "texels": texels, "power": power,
            "rot": (float(np.float32(math.cos(rot))), float(np.float32(math.sin(rot))))}


def build_constant_env(radiance, width=64, height=32):
    """(cdf, pdf) of a constant-radiance map (environment_table)."""
    t = environment_table(radiance, (width, height))
    return t["cdf"], t["texels"][:, 3].copy()


def render(pack, materials, camera_inv_vp, width, height, *, frame_seed, frame_index, max_bounces,
           ray_bias, env_radiance=None, env_size=(64, 32), tile=None, threads=None, events=False,
           sampling=0, light_tree=True, aux=False, emitters=None, env_map=None):
    """One frame (1 spp) of the binary64 integrator mirror.  Returns
    (radiance[h, w, 3] float64, signature[h, w] uint64), plus with `events`
    the decisions mixed into each signature ([h, w, 40] uint64: count, then
    hit keys, light picks, lobes and roulette outcomes in path order), plus
    with `aux` the first-hit images ([h, w, 8]: normal * 0.5 + 0.5, depth,
    base colour, 1)."""
    ps = PackedScene(pack)
    mt = material_table(materials)
    em = np.ascontiguousarray(build_emitters(pack, mt))
    normals = np.ascontiguousarray(pack["normals"], np.float32)
    tri_mat = np.ascontiguousarray(pack["tri_material"], np.uint32)
    env = environment_table(env_radiance, env_size, env_map)
    env_present = env is not None
    if env_present:
        cdf, etex = env["cdf"], env["texels"]
    else:
        cdf, etex = np.zeros(2, np.float32), np.zeros((1, 4), np.float32)
    cdf, etex = np.ascontiguousarray(cdf), np.ascontiguousarray(etex)
    d = RenderDesc()
    d.scene = C.pointer(ps.struct)
    d.normals = normals.ctypes.data_as(F32P)
    d.tri_material = tri_mat.ctypes.data_as(U32P)
    mt = np.ascontiguousarray(mt)
    d.materials = mt.ctypes.data_as(F32P)
    d.material_count = len(materials)
    d.emitters = em.ctypes.data_as(F32P) if em.size else None
    d.emitter_count = em.shape[0]
    d.env_present = 1 if env_present else 0
    d.env_w, d.env_h = ((env["w"], env["h"]) if env_present else (0, 0))
    d.env_cdf = cdf.ctypes.data_as(F32P)
    d.env_texels = etex.ctypes.data_as(F32P)
    d.env_rot = (C.c_float * 2)(*(env["rot"] if env_present else (1.0, 0.0)))
    d.inv_view_proj = (C.c_float * 16)(*np.asarray(camera_inv_vp, np.float32).ravel())
    d.width, d.height = width, height
    d.frame_seed = frame_seed
    d.max_bounces = max_bounces
    d.ray_bias = ray_bias
    d.sampling = sampling
    ana = pack_analytic(emitters or [])
    recs = np.ascontiguousarray(ana[0] if ana[0].size else np.zeros((1, 16), np.float32))
    d.lights = recs.ctypes.data_as(F32P)
    d.n_dir, d.n_point, d.n_spot, d.n_rect = ana[1]
    tree = scene_light_tree(em, env_radiance, env_size, analytic=ana, env_map=env_map) if light_tree else None
    if tree is not None:
        tree = np.ascontiguousarray(tree, np.float32)
        d.light_tree = tree.ctypes.data_as(F32P)
        d.light_tree_nodes = tree.shape[0]
    x0, y0, w, h = tile or (0, 0, width, height)
    out = np.zeros((h, w, 3), np.float64)
    sig = np.zeros((h, w), np.uint64)
    ev = np.zeros((h, w, 40), np.uint64) if events else None
    aux_out = np.zeros((h, w, 8), np.float64) if aux else None
    d.aux = aux_out.ctypes.data_as(F64P) if aux else None
    lib().orc_render(C.byref(d), C.c_uint32(frame_index), x0, y0, w, h,
                     out.ctypes.data_as(F64P), sig.ctypes.data_as(C.POINTER(C.c_uint64)),
                     ev.ctypes.data_as(C.POINTER(C.c_uint64)) if events else None,
                     C.c_uint32(threads or min(16, os.cpu_count() or 1)))
    res = (out, sig) + ((ev,) if events else ()) + ((aux_out,) if aux else ())
    return res


# ---------------------------------------------------------------------------
# Owen-scrambled Sobol streams (sobol.c)
# ---------------------------------------------------------------------------

def _sobol_lib():
    L = lib()
    for name in ("orc_rev32", "orc_sobol_hash"):
        getattr(L, name).restype = C.c_uint32
        getattr(L, name).argtypes = [C.c_uint32]
    for name in ("orc_sobol_combine", "orc_laine_karras", "orc_nested_scramble", "orc_sobol_component",
                 "orc_sobol_rotation", "orc_sobol_frame_key"):
        getattr(L, name).restype = C.c_uint32
        getattr(L, name).argtypes = [C.c_uint32, C.c_uint32]
    L.orc_sobol_owen_u32.restype = C.c_uint32
    L.orc_sobol_owen_u32.argtypes = [C.c_uint32] * 4
    L.orc_sobol_streams.argtypes = [U32P, U32P, C.c_uint32, C.c_uint32, C.c_uint32, U32P]
    L.orc_sobol_stream_state.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, U32P]
    return L


def sobol(name, *args):
    """Call one scalar Sobol primitive of the oracle by name (e.g. 'sobol_hash')."""
    return getattr(_sobol_lib(), "orc_" + name)(*[C.c_uint32(a & 0xFFFFFFFF) for a in args])


def sobol_streams(px, py, key, draws):
    px = np.ascontiguousarray(px, np.uint32)
    py = np.ascontiguousarray(py, np.uint32)
    out = np.zeros((px.size, draws), np.uint32)
    _sobol_lib().orc_sobol_streams(px.ctypes.data_as(U32P), py.ctypes.data_as(U32P), px.size, key, draws,
                                   out.ctypes.data_as(U32P))
    return out


def sobol_stream_state(px, py, key):
    s = np.zeros(7, np.uint32)
    _sobol_lib().orc_sobol_stream_state(px, py, key, s.ctypes.data_as(U32P))
    return dict(zip(["sampleIndex", "dimension", "pixelX", "pixelY", "sequenceKey", "rotationTile",
                     "fallbackState"], [int(x) for x in s]))
