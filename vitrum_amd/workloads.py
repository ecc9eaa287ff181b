"""Deterministic synthetic workloads (scenes, cameras, ray batches).

Every scene is a list of primitive dicts in the shape of vitrum's core `Scene`
primitives (packages/core/src/scene/primitives.ts:167-227): `kind` ('mesh' or
'instanced-mesh'), flat xyz `positions` and `normals` (float32), `indices`
(uint32 triples), an optional column-major `transform` (mesh) or a list of
`instances` (instanced mesh), and a `material` dict.  Generation uses only
numpy with fixed seeds, so the same arguments always give the same bytes.
"""
from __future__ import annotations

import math

import numpy as np

F32 = np.float32


# ---------------------------------------------------------------------------
# Mesh builders
# ---------------------------------------------------------------------------

def _grid_quad(origin, edge_u, edge_v, nu, nv, normal):
    """Tessellate the parallelogram origin + s*edge_u + t*edge_v into nu x nv cells."""
    origin, edge_u, edge_v = (np.asarray(x, np.float64) for x in (origin, edge_u, edge_v))
    s = np.arange(nu + 1) / nu
    t = np.arange(nv + 1) / nv
    ss, tt = np.meshgrid(s, t, indexing="ij")
    pts = origin + ss[..., None] * edge_u + tt[..., None] * edge_v
    pos = pts.reshape(-1, 3)
    i, j = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    a = (i * (nv + 1) + j).ravel()
    b = ((i + 1) * (nv + 1) + j).ravel()
    c, d = b + 1, a + 1
    idx = np.stack([a, b, c, a, c, d], axis=1).reshape(-1, 3).astype(np.uint32)
    # orient triangles to the requested normal
    n = np.cross(edge_u, edge_v)
    if np.dot(n, normal) < 0:
        idx = idx[:, [0, 2, 1]]
    nrm = np.tile(np.asarray(normal, np.float64) / np.linalg.norm(normal), (pos.shape[0], 1))
    return pos, nrm, idx


def _factor_near_square(cells):
    best = (1, cells)
    for a in range(1, int(math.isqrt(cells)) + 1):
        if cells % a == 0:
            best = (a, cells // a)
    return best


def _rot_y(deg):
    r = math.radians(deg)
    c, s = math.cos(r), math.sin(r)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]], np.float64)


def _box_faces(center, size, yaw_deg):
    """Six outward quads of an oriented box: (origin, edge_u, edge_v, normal)."""
    R = _rot_y(yaw_deg)
    hx, hy, hz = (0.5 * s for s in size)
    c = np.asarray(center, np.float64)
    faces = []
    # (corner in local coords, edge_u, edge_v, normal) in local space
    local = [
        ((-hx, -hy, hz), (2 * hx, 0, 0), (0, 2 * hy, 0), (0, 0, 1)),
        ((hx, -hy, -hz), (-2 * hx, 0, 0), (0, 2 * hy, 0), (0, 0, -1)),
        ((hx, -hy, hz), (0, 0, -2 * hz), (0, 2 * hy, 0), (1, 0, 0)),
        ((-hx, -hy, -hz), (0, 0, 2 * hz), (0, 2 * hy, 0), (-1, 0, 0)),
        ((-hx, hy, hz), (2 * hx, 0, 0), (0, 0, -2 * hz), (0, 1, 0)),
        ((-hx, -hy, -hz), (2 * hx, 0, 0), (0, 0, 2 * hz), (0, -1, 0)),
    ]
    for o, u, v, n in local:
        faces.append((c + R @ np.asarray(o), R @ np.asarray(u), R @ np.asarray(v), R @ np.asarray(n)))
    return faces


def _mesh(pos, nrm, idx, material, transform=None):
    prim = {
        "kind": "mesh",
        "positions": np.ascontiguousarray(pos, F32).ravel(),
        "normals": np.ascontiguousarray(nrm, F32).ravel(),
        "indices": np.ascontiguousarray(idx, np.uint32).ravel(),
        "material": material,
    }
    if transform is not None:
        prim["transform"] = np.ascontiguousarray(transform, F32).ravel()
    return prim


def _merge(parts):
    pos, nrm, idx, base = [], [], [], 0
    for p, n, i in parts:
        pos.append(p)
        nrm.append(n)
        idx.append(i + base)
        base += p.shape[0]
    return np.concatenate(pos), np.concatenate(nrm), np.concatenate(idx)


# ---------------------------------------------------------------------------
# Config: Cornell box tessellated to an exact triangle count
# ---------------------------------------------------------------------------

WHITE = {"baseColor": [0.73, 0.73, 0.73], "roughness": 1.0, "metallic": 0.0}
RED = {"baseColor": [0.65, 0.05, 0.05], "roughness": 1.0, "metallic": 0.0}
GREEN = {"baseColor": [0.12, 0.45, 0.15], "roughness": 1.0, "metallic": 0.0}
LIGHT = {"baseColor": [0.0, 0.0, 0.0], "roughness": 1.0, "metallic": 0.0,
         "emissive": [1.0, 0.85, 0.6], "emissiveIntensity": 17.0}


def cornell_scene(total_triangles: int = 32768):
    """Cornell-style box: 5 walls, 2 rotated boxes, 1 quad light.

    Each object is one mesh primitive.  Every face is a uniform grid; cell
    counts are apportioned by face area and the largest face absorbs the
    remainder, so the scene has exactly `total_triangles` triangles.
    """
    if total_triangles % 2:
        raise ValueError("total_triangles must be even")
    walls = [  # origin, edge_u, edge_v, inward normal, material
        ((-1, 0, -1), (2, 0, 0), (0, 0, 2), (0, 1, 0), WHITE),      # floor
        ((-1, 2, -1), (2, 0, 0), (0, 0, 2), (0, -1, 0), WHITE),     # ceiling
        ((-1, 0, -1), (2, 0, 0), (0, 2, 0), (0, 0, 1), WHITE),      # back
        ((-1, 0, -1), (0, 0, 2), (0, 2, 0), (1, 0, 0), RED),        # left
        ((1, 0, -1), (0, 0, 2), (0, 2, 0), (-1, 0, 0), GREEN),      # right
    ]
    light = ((-0.25, 1.98, -0.25), (0.5, 0, 0), (0, 0, 0.5), (0, -1, 0), LIGHT)
    tall = _box_faces((-0.33, 0.6, -0.28), (0.56, 1.2, 0.56), 15.0)
    short = _box_faces((0.34, 0.3, 0.3), (0.56, 0.6, 0.56), -18.0)
    faces = [(w[0], w[1], w[2], w[3]) for w in walls] + [light[:4]] + tall + short
    owner = list(range(5)) + [5] + [6] * 6 + [7] * 6
    areas = np.array([np.linalg.norm(np.cross(f[1], f[2])) for f in faces])
    cells_total = total_triangles // 2
    cells = np.maximum(1, np.floor(areas / areas.sum() * cells_total)).astype(int)
    cells[int(np.argmax(areas))] += cells_total - cells.sum()
    parts = {k: [] for k in range(8)}
    for f, n_cells, o in zip(faces, cells, owner):
        a, b = _factor_near_square(int(n_cells))
        lu, lv = np.linalg.norm(f[1]), np.linalg.norm(f[2])
        nu, nv = (b, a) if lu >= lv else (a, b)
        parts[o].append(_grid_quad(f[0], f[1], f[2], nu, nv, f[3]))
    mats = [WHITE, WHITE, WHITE, RED, GREEN, LIGHT, WHITE, WHITE]
    prims = []
    for k in range(8):
        p, n, i = _merge(parts[k])
        prims.append(_mesh(p, n, i, mats[k]))
    assert sum(p["indices"].size // 3 for p in prims) == total_triangles
    camera = look_at_camera(eye=(0.0, 1.0, 2.7), target=(0.0, 1.0, 0.0), up=(0, 1, 0), vfov_deg=60.0)
    return {"primitives": prims, "camera": camera, "scene_radius": math.sqrt(3.0),
            "sky": [0.0, 0.0, 0.0]}


# ---------------------------------------------------------------------------
# Config: instanced displaced spheres
# ---------------------------------------------------------------------------

def octa_sphere(n: int = 16, split: bool = True):
    """Octahedron with each face subdivided n x n, projected to the unit sphere.

    8 * n^2 triangles; with split=True each triangle is halved through the
    midpoint of its first edge, giving 16 * n^2 (n=16: 4,096 triangles).
    """
    corners = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, 0, 0], [0, -1, 0], [0, 0, -1]], float)
    faces = [(0, 1, 2), (1, 3, 2), (3, 4, 2), (4, 0, 2), (1, 0, 5), (3, 1, 5), (4, 3, 5), (0, 4, 5)]
    key_to_index: dict = {}
    verts: list = []

    def vid(p):
        q = p / np.linalg.norm(p)
        key = tuple(np.round(q * 1e9).astype(np.int64))
        if key not in key_to_index:
            key_to_index[key] = len(verts)
            verts.append(q)
        return key_to_index[key]

    tris = []
    for a, b, c in faces:
        A, B, Cc = corners[a], corners[b], corners[c]
        grid = {}
        for i in range(n + 1):
            for j in range(n + 1 - i):
                grid[(i, j)] = A + (B - A) * (i / n) + (Cc - A) * (j / n)
        for i in range(n):
            for j in range(n - i):
                tris.append((grid[(i, j)], grid[(i + 1, j)], grid[(i, j + 1)]))
                if j < n - i - 1:
                    tris.append((grid[(i + 1, j)], grid[(i + 1, j + 1)], grid[(i, j + 1)]))
    out = []
    for a, b, c in tris:
        if split:
            m = 0.5 * (a + b)
            out.append((vid(a), vid(m), vid(c)))
            out.append((vid(m), vid(b), vid(c)))
        else:
            out.append((vid(a), vid(b), vid(c)))
    return np.asarray(verts, np.float64), np.asarray(out, np.uint32)


def _displace(verts, rng, amplitude=0.1, waves=6):
    """Radial displacement by a seeded sum of smooth waves, |d| <= amplitude."""
    dirs = rng.normal(size=(waves, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    freq = rng.uniform(2.0, 6.0, size=waves)
    phase = rng.uniform(0, 2 * np.pi, size=waves)
    w = np.sin(verts @ dirs.T * freq + phase).sum(axis=1) / waves
    return verts * (1.0 + amplitude * w)[:, None]


def _vertex_normals(pos, idx):
    tri = pos[idx]
    fn = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    vn = np.zeros_like(pos)
    for k in range(3):
        np.add.at(vn, idx[:, k], fn)
    ln = np.linalg.norm(vn, axis=1, keepdims=True)
    ln[ln == 0] = 1
    return vn / ln


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def _mat4(R, t):
    m = np.eye(4)
    m[:3, :3] = R
    m[:3, 3] = t
    return m.T.ravel()  # column-major


def spheres_scene(instances: int = 256, seed: int = 2024, radius: float = 4.0,
                  lights: int = 8, box: float = 100.0, sphere_subdiv: int = 16):
    """Instances of displaced octa-spheres (4,096 triangles each) with random
    rigid transforms uniform in a box^3 volume, plus emissive light spheres.

    Each displaced sphere is its own mesh (distinct displacement seed), so the
    scene holds instances * 4,096 unique triangles.  Materials: Lambert albedo
    U[0.1, 0.9]^3 and GGX roughness U[0.05, 1.0].
    """
    rng = np.random.default_rng(seed)
    base_v, base_i = octa_sphere(sphere_subdiv, split=True)
    prims = []
    for k in range(instances):
        v = _displace(base_v, rng) * radius
        n = _vertex_normals(v, base_i)
        mat = {"baseColor": rng.uniform(0.1, 0.9, 3).tolist(),
               "roughness": float(rng.uniform(0.05, 1.0)), "metallic": 0.0}
        xf = _mat4(_random_rotation(rng), rng.uniform(0.0, box, 3))
        prims.append(_mesh(v, n, base_i, mat, xf))
    light_v, light_i = octa_sphere(8, split=False)
    for k in range(lights):
        v = light_v * (0.5 * radius)
        n = _vertex_normals(v, light_i)
        col = rng.uniform(0.6, 1.0, 3)
        mat = {"baseColor": [0.0, 0.0, 0.0], "roughness": 1.0, "metallic": 0.0,
               "emissive": (col / col.max()).tolist(), "emissiveIntensity": 40.0}
        xf = _mat4(np.eye(3), rng.uniform(0.0, box, 3))
        prims.append(_mesh(v, n, light_i, mat, xf))
    c = box * 0.5
    camera = look_at_camera(eye=(c, c, -0.45 * box), target=(c, c, c), up=(0, 1, 0), vfov_deg=60.0)
    return {"primitives": prims, "camera": camera, "scene_radius": box * math.sqrt(3) * 0.5 + radius,
            "sky": [0.3, 0.35, 0.45]}


def finned_latlong_sphere(segments: int = 61, rings: int = 21):
    """Unit lat-long sphere (pole fans plus quad bands: 2 * segments *
    (rings - 1) triangles; 61 x 21 gives 2,440) with one extra fin triangle
    standing on the equator, so the mesh has exactly 2,441 triangles.  A closed
    two-manifold triangle mesh always has an even triangle count, so an odd
    count needs an open or non-manifold part; the fin is that part."""
    verts = [[0.0, 1.0, 0.0]]
    for r in range(1, rings):
        th = np.pi * r / rings
        for s_ in range(segments):
            ph = 2 * np.pi * s_ / segments
            verts.append([np.sin(th) * np.cos(ph), np.cos(th), np.sin(th) * np.sin(ph)])
    verts.append([0.0, -1.0, 0.0])
    bottom = len(verts) - 1

    def ring_v(r, s_):
        return 1 + (r - 1) * segments + (s_ % segments)

    tris = []
    for s_ in range(segments):
        tris.append((0, ring_v(1, s_ + 1), ring_v(1, s_)))
    for r in range(1, rings - 1):
        for s_ in range(segments):
            a, b = ring_v(r, s_), ring_v(r, s_ + 1)
            c, d = ring_v(r + 1, s_ + 1), ring_v(r + 1, s_)
            tris += [(a, b, c), (a, c, d)]
    for s_ in range(segments):
        tris.append((bottom, ring_v(rings - 1, s_), ring_v(rings - 1, s_ + 1)))
    eq = rings // 2
    a, b = ring_v(eq, 0), ring_v(eq, 1)
    verts.append([1.3, 0.0, 0.05])
    tris.append((a, b, len(verts) - 1))
    return np.asarray(verts, np.float64), np.asarray(tris, np.uint32)


def large_scene(instances: int = 4096, seed: int = 77, radius: float = 4.0, lights: int = 32,
                box: float = 252.0):
    """Config-5-sized scene: `instances` distinct displaced 2,441-triangle
    meshes (finned_latlong_sphere; 4,096 x 2,441 = 9,998,336 triangles) with
    random rigid transforms uniform in a box^3 volume (box chosen so the
    instance density matches the 256-instance scene's 100^3), emissive light
    spheres and a constant sky.  Materials as in spheres_scene."""
    rng = np.random.default_rng(seed)
    base_v, base_i = finned_latlong_sphere()
    prims = []
    for k in range(instances):
        v = _displace(base_v, rng) * radius
        n = _vertex_normals(v, base_i)
        mat = {"baseColor": rng.uniform(0.1, 0.9, 3).tolist(),
               "roughness": float(rng.uniform(0.05, 1.0)), "metallic": 0.0}
        xf = _mat4(_random_rotation(rng), rng.uniform(0.0, box, 3))
        prims.append(_mesh(v, n, base_i, mat, xf))
    light_v, light_i = octa_sphere(8, split=False)
    for k in range(lights):
        v = light_v * (0.5 * radius)
        n = _vertex_normals(v, light_i)
        col = rng.uniform(0.6, 1.0, 3)
        mat = {"baseColor": [0.0, 0.0, 0.0], "roughness": 1.0, "metallic": 0.0,
               "emissive": (col / col.max()).tolist(), "emissiveIntensity": 40.0}
        xf = _mat4(np.eye(3), rng.uniform(0.0, box, 3))
        prims.append(_mesh(v, n, light_i, mat, xf))
    c = box * 0.5
    camera = look_at_camera(eye=(c, c, -0.45 * box), target=(c, c, c), up=(0, 1, 0), vfov_deg=60.0)
    return {"primitives": prims, "camera": camera, "scene_radius": box * math.sqrt(3) * 0.5 + radius,
            "sky": [0.3, 0.35, 0.45]}


# ---------------------------------------------------------------------------
# Cameras and rays
# ---------------------------------------------------------------------------

def look_at_camera(eye, target, up, vfov_deg, near=0.01, far=1000.0):
    eye, target, up = (np.asarray(x, np.float64) for x in (eye, target, up))
    f = target - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    view = np.eye(4)
    view[0, :3], view[1, :3], view[2, :3] = s, u, -f
    view[:3, 3] = -view[:3, :3] @ eye
    return {"eye": eye, "view": view, "vfov_deg": vfov_deg, "near": near, "far": far}


def projection(vfov_deg, aspect, near, far):
    fy = 1.0 / math.tan(math.radians(vfov_deg) * 0.5)
    p = np.zeros((4, 4))
    p[0, 0] = fy / aspect
    p[1, 1] = fy
    p[2, 2] = (far + near) / (near - far)
    p[2, 3] = 2 * far * near / (near - far)
    p[3, 2] = -1
    return p


def inv_view_proj(camera, width, height):
    """Column-major float32 inverse of projection x view (FrameInput's
    viewMatrix / projMatrix, packages/core/src/frame.ts:91)."""
    proj = projection(camera["vfov_deg"], width / height, camera["near"], camera["far"])
    inv = np.linalg.inv(proj @ camera["view"])
    return inv.T.astype(np.float32).ravel()


RAY_DTYPE = np.dtype([("o", "<f4", 3), ("tmin", "<f4"), ("d", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("prim", "<u4"), ("inst", "<u4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4")])
TMAX_INFINITY = np.uint32(0x7F7FFFFF).view(np.float32)


def pcg_init(px, py, seed):
    """pcgInit (packages/shared-samplers/src/wgsl/pcg.wgsl.ts:5-11), vectorised."""
    px = np.asarray(px, np.uint32)
    py = np.asarray(py, np.uint32)
    with np.errstate(over="ignore"):
        s = px * np.uint32(1664525) + py * np.uint32(1013904223) + np.uint32(seed) * np.uint32(22695477)
        s ^= s >> np.uint32(17)
        s ^= s << np.uint32(31)
        s ^= s >> np.uint32(11)
    return s.astype(np.uint32)


def pcg_next(state):
    """pcgNext (pcg.wgsl.ts:13-18); returns (new_state, word)."""
    with np.errstate(over="ignore"):
        s = (state * np.uint32(747796405) + np.uint32(2891336453)).astype(np.uint32)
        word = ((s >> ((s >> np.uint32(28)) + np.uint32(4))) ^ s) * np.uint32(277803737)
        word = (word >> np.uint32(22)) ^ word
    return s, word.astype(np.uint32)


def rand_f32(state):
    state, w = pcg_next(state)
    return state, (w >> np.uint32(8)).astype(F32) / F32(16777216.0)


def ray_bias(scene_radius):
    """ptWebgpuRayOriginBias / ptWebgpuRayTMin (pt-webgpu/src/scene/sceneScalePolicy.ts:83-123)."""
    bias = F32(scene_radius * 1e-3 / math.sqrt(3.0))
    return bias, F32(bias * F32(0.1))


def primary_rays(camera, width, height, frame_seed=0x9E3779B9, frame_index=0, scene_radius=1.0):
    """Pinhole rays with PCG sub-pixel jitter through the near-plane unprojection
    of generatePrimaryRay / unproject_ray_common (pt-webgpu/src/wgsl/
    pathTrace/kernelCore.wgsl.ts:22-26, common.wgsl.ts:431-469).

    Returns (rays, rng_state) with one ray per pixel in row-major order.
    """
    px, py = np.meshgrid(np.arange(width, dtype=np.uint32), np.arange(height, dtype=np.uint32))
    px, py = px.ravel(), py.ravel()
    st = pcg_init(px, py, np.uint32(frame_seed) ^ np.uint32(frame_index))
    st, jx = rand_f32(st)
    st, jy = rand_f32(st)
    u = (px.astype(F32) + jx) / F32(width)
    v = (py.astype(F32) + jy) / F32(height)
    ndc = np.stack([u * F32(2) - F32(1), F32(1) - v * F32(2)], axis=1).astype(np.float64)
    proj = projection(camera["vfov_deg"], width / height, camera["near"], camera["far"])
    inv = np.linalg.inv(proj @ camera["view"])
    ones = np.ones(ndc.shape[0])
    far = inv @ np.stack([ndc[:, 0], ndc[:, 1], ones, ones])
    near = inv @ np.stack([ndc[:, 0], ndc[:, 1], -ones, ones])
    far = (far[:3] / far[3]).T
    near = (near[:3] / near[3]).T
    d = far - near
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros(px.size, RAY_DTYPE)
    rays["o"] = near.astype(F32)
    rays["d"] = _unit_f32(d.astype(F32))
    _, tmin = ray_bias(scene_radius)
    rays["tmin"] = tmin
    rays["tmax"] = TMAX_INFINITY
    return rays, st


def _unit_f32(d):
    d = np.asarray(d, F32)
    s = np.max(np.abs(d), axis=1, keepdims=True)
    s[s == 0] = 1
    q = d / s
    ln = np.sqrt(np.sum(q * q, axis=1, keepdims=True, dtype=F32)).astype(F32)
    return (q / ln).astype(F32)


def world_triangles(pack, prim, inst):
    """World-space f32 vertices of hit triangles (for bounce-ray synthesis)."""
    idx = pack["indices"].reshape(-1, 4)[prim, :3]
    pos = pack["positions"].reshape(-1, 4)[:, :3].astype(np.float64)
    l2w = pack["tlas_l2w"].reshape(-1, 4, 4).astype(np.float64)  # [inst][col][row]
    M = np.transpose(l2w[inst], (0, 2, 1))  # row-major
    out = []
    for k in range(3):
        p = np.concatenate([pos[idx[:, k]], np.ones((prim.size, 1))], axis=1)
        w = np.einsum("nij,nj->ni", M, p)
        out.append(w[:, :3] / w[:, 3:4])
    return out


def bounce_rays(pack, rays, hits, rng_state, scene_radius):
    """One cosine-hemisphere bounce ray per primary hit (buildOnb /
    cosineHemisphereSample, pt-webgpu/src/wgsl/pathTrace/bsdf.wgsl.ts:1632-1664),
    origin offset along the geometric normal by the ray-origin bias."""
    hit = hits["inst"] != 0xFFFFFFFF
    sel = np.nonzero(hit)[0]
    st = rng_state[sel]
    st, u1 = rand_f32(st)
    st, u2 = rand_f32(st)
    a, b, c = world_triangles(pack, hits["prim"][sel], hits["inst"][sel])
    n = np.cross(b - a, c - a)
    n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)
    d = rays["d"][sel].astype(np.float64)
    flip = np.sum(n * d, axis=1) > 0
    n[flip] *= -1
    p = rays["o"][sel].astype(np.float64) + d * hits["t"][sel, None].astype(np.float64)
    bias, tmin = ray_bias(scene_radius)
    origin = p + n * float(bias)
    r = np.sqrt(u1.astype(np.float64))
    phi = 2 * np.pi * u2.astype(np.float64)
    cos_t = np.sqrt(np.maximum(0.0, 1.0 - u1.astype(np.float64)))
    up = np.where(np.abs(n[:, 1:2]) > 0.999, [[1.0, 0, 0]], [[0, 1.0, 0]])
    t = np.cross(up, n)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    bt = np.cross(n, t)
    wi = (r * np.cos(phi))[:, None] * t + (r * np.sin(phi))[:, None] * bt + cos_t[:, None] * n
    out = np.zeros(sel.size, RAY_DTYPE)
    out["o"] = origin.astype(F32)
    out["d"] = _unit_f32(wi.astype(F32))
    out["tmin"] = tmin
    out["tmax"] = TMAX_INFINITY
    return out


def batch_workload(count: int = 64, seed: int = 3, lo: int = 1_000, hi: int = 2_000_000):
    """The batched-launch workload: `count` Cornell-style scenes whose
    triangle counts are log-uniform in [lo, hi] (rounded to even; each scene
    tessellated as cornell_scene), each with a camera resolution drawn from
    {256^2, 512^2, 1024^2}.  Returns [(scene, resolution)]."""
    rng = np.random.default_rng(seed)
    tris = np.exp(rng.uniform(np.log(lo), np.log(hi), count))
    res = rng.choice([256, 512, 1024], count)
    out = []
    for t, r in zip(tris, res):
        n = max(64, int(round(t / 2)) * 2)
        out.append((cornell_scene(n), int(r)))
    return out


def config_scene(name: str):
    if name == "cornell":
        return cornell_scene()
    if name == "spheres":
        return spheres_scene()
    raise KeyError(name)
