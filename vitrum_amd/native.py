"""ctypes binding of libvthip.so (the C ABI declared in include/vitrum_hip.h).

Used by the Python tests and the benchmark.  Loading fails loudly when the
library has not been built: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .workloads import HIT_DTYPE, RAY_DTYPE

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvthip.so")
_LIB = None

U32P = C.POINTER(C.c_uint32)
F32P = C.POINTER(C.c_float)
F64P = C.POINTER(C.c_double)

STATUS = {0: "VT_OK", -1: "VT_ERR_INVALID_ARGUMENT", -2: "VT_ERR_RANGE", -3: "VT_ERR_BUILD",
          -4: "VT_ERR_HIP", -5: "VT_ERR_NO_DEVICE", -6: "VT_ERR_STATE", -7: "VT_ERR_OUT_OF_MEMORY"}

EXPORTED = [
    "vt_create", "vt_destroy", "vt_last_error", "vt_set_scene", "vt_trace_closest", "vt_trace_any",
    "vt_trace_closest_device", "vt_trace_any_device", "vt_set_counting", "vt_trace_closest_counts",
    "vt_set_trace_tuning", "vt_selftest_fp32", "vt_device_alloc", "vt_device_free",
    "vt_memcpy_h2d", "vt_memcpy_d2h", "vt_memcpy_d2d", "vt_synchronize", "vt_get_stream", "vt_stats",
    "vt_build_blas", "vt_build_tlas", "vt_pack_scene", "vt_pack_field_words",
    "vt_pack_copy_field", "vt_pack_free", "vt_version", "vt_render_frame", "vt_capture",
    "vt_reset", "vt_frame_time", "vt_capture_signatures", "vt_render_info", "vt_light_tree", "vt_invert_mat4",
    "vt_set_bvh_traversal", "vt_set_node_width", "vt_set_hit_filter",
    "vt_nrc_param_count", "vt_nrc_reset", "vt_nrc_set_weights", "vt_nrc_get_weights", "vt_nrc_infer",
    "vt_nrc_infer_device", "vt_nrc_train_step", "vt_nrc_train_step_device", "vt_nrc_last_loss",
    "vt_nrc_gradient", "vt_nrc_render_config", "vt_nrc_teacher_vertices", "vt_nrc_timing", "vt_batch_create", "vt_batch_destroy",
    "vt_batch_add_scene", "vt_batch_scene_count", "vt_batch_trace_closest_device", "vt_batch_capture",
    "vt_batch_replay", "vt_batch_autotune", "vt_batch_set_tuning",
    "vt_refit_bvh_bounds_device", "vt_refit_primitive", "vt_scene_copy_field", "vt_build_blas_device",
    "vt_set_device_build",
]

SCENE_BLAS_NODE64, SCENE_BLAS_NODE128, SCENE_TRIANGLES = 64, 65, 66
PACK_FIELDS = ["positions", "normals", "indices", "tri_material", "tri_source", "tri_primitive",
               "bvh_nodes", "tlas_nodes", "tlas_instance_indices", "tlas_blas_roots", "tlas_w2l",
               "tlas_l2w", "instance_primitive", "instance_source"]
FLOAT_PACK_FIELDS = {"positions", "normals", "tlas_w2l", "tlas_l2w"}


class Tuning(C.Structure):
    _fields_ = [("refill_min", C.c_uint32), ("steps", C.c_uint32), ("near_first", C.c_uint32),
                ("defer_min", C.c_uint32), ("blocks_per_cu", C.c_uint32), ("ms", C.c_float)]


class NrcOptions(C.Structure):
    _fields_ = [("learning_rate", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float),
                ("epsilon", C.c_float), ("seed", C.c_uint32), ("table_learning_rate", C.c_float),
                ("reserved", C.c_uint32 * 2)]


NRC_FEATURES = 16


class VtError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


class Primitive(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("material", C.c_uint32), ("positions", F32P),
                ("position_floats", C.c_uint32), ("normals", F32P), ("normal_floats", C.c_uint32),
                ("indices", U32P), ("index_count", C.c_uint32), ("transforms", F32P),
                ("transform_count", C.c_uint32)]


class Material(C.Structure):
    _fields_ = [("base_color", C.c_float * 3), ("roughness", C.c_float), ("metallic", C.c_float),
                ("emissive", C.c_float * 3), ("emissive_intensity", C.c_float),
                ("double_sided", C.c_uint32), ("clearcoat", C.c_float), ("clearcoat_roughness", C.c_float),
                ("sheen", C.c_float), ("sheen_roughness", C.c_float), ("sheen_color", C.c_float * 3),
                ("anisotropy", C.c_float), ("anisotropy_rotation", C.c_float), ("has_specular", C.c_uint32),
                ("specular_color", C.c_float * 3), ("specular_intensity", C.c_float)]


class Environment(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("radiance", C.c_float * 3), ("texels", F32P),
                ("width", C.c_uint32), ("height", C.c_uint32), ("channels", C.c_uint32),
                ("intensity", C.c_float), ("rotation_y", C.c_float), ("sun_direction", C.c_double * 3),
                ("turbidity", C.c_double), ("rayleigh", C.c_double), ("mie_coefficient", C.c_double),
                ("mie_directional_g", C.c_double)]


def _sky_environment(sky) -> "Environment":
    """ProceduralSkyEnvironment (environment.ts:59-71) with the reference's
    defaults (environmentPacking.ts:245-262)."""
    env = Environment()
    env.kind = 3
    env.sun_direction = (C.c_double * 3)(*sky.get("sunDirection", (0.0, 1.0, 0.0)))
    env.turbidity = float(sky.get("turbidity", 2.0))
    env.rayleigh = float(sky.get("rayleigh", 1.0))
    env.mie_coefficient = float(sky.get("mieCoefficient", 0.005))
    env.mie_directional_g = float(sky.get("mieDirectionalG", 0.8))
    env.intensity = float(sky.get("intensity", 1.0))
    return env


def bake_procedural_sky(sky, width=256, height=128):
    """vt_bake_procedural_sky: the Preetham map ((height, width, 3) float32)
    and its luminance x solid-angle integral; host only, no device."""
    env = _sky_environment(sky)
    out = np.zeros((height, width, 3), np.float32)
    integral = C.c_double(0.0)
    check(lib().vt_bake_procedural_sky(C.byref(env), width, height, out.ctypes.data_as(F32P), C.byref(integral)))
    return out, integral.value


def _environment(scene):
    """(Environment, keep-alive) of a scene dict: "hdri" = {"texels": (H, W, C)
    float32 array, "intensity", "rotation_y"} (HdriEnvironment), else "sky" =
    constant radiance [r, g, b]."""
    if scene.get("procedural_sky") is not None:
        return _sky_environment(scene["procedural_sky"]), None
    hdri = scene.get("hdri")
    if hdri is not None:
        tex = np.ascontiguousarray(hdri["texels"], dtype=np.float32)
        if tex.ndim == 2:
            tex = tex[:, :, None]
        h, w, ch = tex.shape
        env = Environment(2, (C.c_float * 3)(0, 0, 0), tex.ctypes.data_as(F32P), w, h, ch,
                          float(hdri.get("intensity", 1.0)), float(hdri.get("rotation_y", 0.0)))
        return env, tex
    sky = scene.get("sky", [0.0, 0.0, 0.0])
    return Environment(1 if any(sky) else 0, (C.c_float * 3)(*sky)), None


class Emitter(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("color", C.c_float * 3), ("intensity", C.c_float),
                ("cast_shadow", C.c_uint32), ("position", C.c_float * 3), ("direction", C.c_float * 3),
                ("normal", C.c_float * 3), ("u_axis", C.c_float * 3), ("v_axis", C.c_float * 3),
                ("radius", C.c_float), ("angle", C.c_float), ("penumbra", C.c_float), ("distance", C.c_float),
                ("decay", C.c_float), ("angular_diameter", C.c_float)]


EMITTER_KINDS = {"directional": 0, "point": 1, "spot": 2, "rect-area": 3, "disc-area": 4}


def _emitter(e) -> Emitter:
    """A SceneEmitter dict (packages/core/src/scene/emitters.ts field names)."""
    v3 = lambda k: (C.c_float * 3)(*e.get(k, (0.0, 0.0, 0.0)))  # noqa: E731
    return Emitter(EMITTER_KINDS[e["kind"]], v3("color"), float(e.get("intensity", 1.0)),
                   0 if e.get("castShadow", True) is False else 1, v3("position"), v3("direction"), v3("normal"),
                   v3("uAxis"), v3("vAxis"), float(e.get("radius", 0.0)), float(e.get("angle", 0.0)),
                   float(e.get("penumbra", 0.0)), float(e.get("distance", 0.0)), float(e.get("decay", 2.0)),
                   float(e.get("angularDiameter", 0.0)))


class SceneDesc(C.Structure):
    _fields_ = [("primitives", C.POINTER(Primitive)), ("primitive_count", C.c_uint32),
                ("materials", C.POINTER(Material)), ("material_count", C.c_uint32),
                ("environment", Environment), ("scene_radius", C.c_float),
                ("emitters", C.POINTER(Emitter)), ("emitter_count", C.c_uint32)]


class Options(C.Structure):
    _fields_ = [("max_bounces", C.c_uint32), ("max_samples_per_pixel", C.c_uint32),
                ("sampling", C.c_uint32), ("light_selection", C.c_uint32), ("bvh_traversal", C.c_uint32),
                ("reserved", C.c_uint32 * 3)]


class Present(C.Structure):
    _fields_ = [("tonemap", C.c_uint32), ("exposure", C.c_float), ("color_space", C.c_uint32)]


class Frame(C.Structure):
    _fields_ = [("inv_view_proj", C.c_float * 16), ("width", C.c_uint32), ("height", C.c_uint32),
                ("frame_index", C.c_uint32), ("frame_seed", C.c_uint32),
                ("max_bounces", C.c_uint32), ("samples_target", C.c_uint32)]


class FrameStats(C.Structure):
    _fields_ = [("last_trace_ms", C.c_float), ("last_frame_ms", C.c_float),
                ("samples_accumulated", C.c_uint32), ("last_ray_count", C.c_uint32),
                ("device_bytes", C.c_uint64)]


PACK_FIELDS = ["positions", "normals", "indices", "tri_material", "tri_source", "tri_primitive",
               "bvh_nodes", "tlas_nodes", "tlas_instance_indices", "tlas_blas_roots", "tlas_w2l",
               "tlas_l2w", "instance_primitive", "instance_source"]
FLOAT_FIELDS = {"positions", "normals", "tlas_w2l", "tlas_l2w"}


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() or "
                              "make -C vitrum_amd/csrc")
        L = C.CDLL(LIB_PATH)
        L.vt_version.restype = C.c_char_p
        L.vt_pack_field_words.restype = C.c_int64
        L.vt_pack_field_words.argtypes = [C.c_void_p, C.c_int32]
        L.vt_pack_copy_field.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_uint64]
        L.vt_pack_free.argtypes = [C.c_void_p]
        L.vt_device_alloc.argtypes = [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)]
        L.vt_refit_bvh_bounds_device.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32,
                                                 C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32]
        L.vt_refit_primitive.argtypes = [C.c_void_p, C.c_uint32, F32P, C.c_uint32, F32P, C.c_uint32]
        L.vt_scene_copy_field.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_uint64,
                                          C.POINTER(C.c_uint64)]
        L.vt_device_free.argtypes = [C.c_void_p, C.c_void_p]
        L.vt_memcpy_h2d.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
        L.vt_memcpy_d2h.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
        L.vt_memcpy_d2d.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
        for name in ("vt_trace_closest", "vt_trace_any", "vt_trace_closest_device",
                     "vt_trace_any_device"):
            getattr(L, name).argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]
        L.vt_trace_closest_counts.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p,
                                              C.c_void_p]
        L.vt_capture_output.argtypes = [C.c_void_p, F32P, C.c_uint32, C.c_uint32, C.c_void_p]
        L.vt_capture_aux.argtypes = [C.c_void_p, C.c_uint32, F32P, C.c_uint32, C.c_uint32]
        L.vt_bake_procedural_sky.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, F32P, C.POINTER(C.c_double)]
        L.vt_render_batch_create.argtypes = [C.c_void_p, C.c_void_p]
        L.vt_render_batch_destroy.argtypes = [C.c_void_p]
        L.vt_render_batch_add.argtypes = [C.c_void_p, C.c_void_p]
        L.vt_render_batch_frame.argtypes = [C.c_void_p, C.c_void_p]
        L.vt_render_batch_capture.argtypes = [C.c_void_p, C.c_void_p]
        L.vt_render_batch_replay.argtypes = [C.c_void_p, C.c_uint32]
        _LIB = L
    return _LIB


def last_error(eng=None) -> str:
    buf = C.create_string_buffer(1024)
    lib().vt_last_error(eng, buf, C.c_size_t(len(buf)))
    return buf.value.decode(errors="replace")


def check(status, eng=None):
    if status != 0:
        raise VtError(status, last_error(eng))


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def _prim_structs(primitives):
    keep = []
    arr = (Primitive * max(1, len(primitives)))()
    for i, p in enumerate(primitives):
        pos = np.ascontiguousarray(p["positions"], np.float32).ravel()
        nrm = np.ascontiguousarray(p["normals"], np.float32).ravel()
        idx = p.get("indices")
        idx = None if idx is None else np.ascontiguousarray(idx, np.uint32).ravel()
        if p["kind"] == "instanced-mesh":
            xf = np.ascontiguousarray(np.asarray(p["instances"], np.float32).reshape(-1, 16))
            kind, ntx = 1, xf.shape[0]
        else:
            t = p.get("transform")
            xf = None if t is None else np.ascontiguousarray(np.asarray(t, np.float32).reshape(16))
            kind, ntx = 0, (0 if xf is None else 1)
        keep += [pos, nrm, idx, xf]
        arr[i] = Primitive(kind, p.get("material_id", i), _p(pos, F32P), pos.size, _p(nrm, F32P),
                           nrm.size, _p(idx, U32P), 0 if idx is None else idx.size, _p(xf, F32P), ntx)
    return arr, keep


def inject_failure(flags):
    """Test hook without an engine: bit 1 makes the next vt_pack_scene on this
    thread throw std::bad_alloc inside the library."""
    check(lib().vt_debug_inject_failure(None, C.c_uint32(flags)))


def pack_scene(primitives, threads=0):
    """packSceneFromCore through the product builder (vt_pack_scene)."""
    arr, _keep = _prim_structs(primitives)
    handle = C.c_void_p()
    check(lib().vt_pack_scene(arr, C.c_uint32(len(primitives)), C.c_uint32(threads), C.byref(handle)))
    out = {}
    try:
        for fid, name in enumerate(PACK_FIELDS):
            n = lib().vt_pack_field_words(handle, fid)
            buf = np.zeros(max(n, 1), np.float32 if name in FLOAT_FIELDS else np.uint32)
            check(lib().vt_pack_copy_field(handle, fid, buf.ctypes.data, C.c_uint64(buf.size)))
            out[name] = buf[:n]
    finally:
        lib().vt_pack_free(handle)
    return out


def build_blas(positions, indices, tri_material_ids=None, *, position_stride=4, index_stride=4,
               max_leaf=4, num_bins=16, max_depth=58, engine=None):
    """buildArrayBvh through the product builder: vt_build_blas on the host,
    or vt_build_blas_device on `engine`'s GPU when an Engine is given."""
    positions = np.ascontiguousarray(positions, np.float32).ravel()
    indices = np.ascontiguousarray(indices, np.uint32).ravel()
    n = indices.size // index_stride
    mats = np.zeros(n, np.uint32) if tri_material_ids is None else np.ascontiguousarray(
        tri_material_ids, np.uint32)
    nodes = np.zeros(max(1, 2 * n - 1) * 8, np.uint32)
    count = C.c_uint32()
    ridx = np.zeros(max(indices.size, 1), np.uint32)
    rmat = np.zeros(max(n, 1), np.uint32)
    rsrc = np.zeros(max(n, 1), np.uint32)
    args = (_p(positions, F32P), C.c_uint32(positions.size), C.c_uint32(position_stride), _p(indices, U32P),
            C.c_uint32(indices.size), C.c_uint32(index_stride), _p(mats, U32P), C.c_uint32(max_leaf),
            C.c_uint32(num_bins), C.c_uint32(max_depth), _p(nodes, U32P), C.byref(count), _p(ridx, U32P),
            _p(rmat, U32P), _p(rsrc, U32P))
    if engine is None:
        check(lib().vt_build_blas(*args))
    else:
        check(lib().vt_build_blas_device(engine.handle, *args), engine.handle)
    c = count.value
    return {"nodes": nodes[: c * 8].copy(), "node_count": c,
            "reordered_indices": ridx[: indices.size].copy(), "reordered_mat": rmat[:n].copy(),
            "reordered_to_source": rsrc[:n].copy()}


def build_tlas(blas_ids, aabb_min, aabb_max, w2l, *, max_leaf=1, num_bins=16, max_depth=58):
    blas_ids = np.ascontiguousarray(blas_ids, np.uint32)
    n = blas_ids.size
    mn = np.ascontiguousarray(aabb_min, np.float64).ravel()
    mx = np.ascontiguousarray(aabb_max, np.float64).ravel()
    xf = np.ascontiguousarray(w2l, np.float32).ravel()
    nodes = np.zeros(max(1, 2 * n - 1) * 8, np.uint32)
    count = C.c_uint32()
    perm = np.zeros(max(n, 1), np.uint32)
    roots = np.zeros(max(n, 1), np.uint32)
    tf = np.zeros(max(n, 1) * 16, np.float32)
    check(lib().vt_build_tlas(C.c_uint32(n), _p(blas_ids, U32P), _p(mn, F64P), _p(mx, F64P),
                              _p(xf, F32P), C.c_uint32(max_leaf), C.c_uint32(num_bins),
                              C.c_uint32(max_depth), _p(nodes, U32P), C.byref(count),
                              _p(perm, U32P), _p(roots, U32P), _p(tf, F32P)))
    c = count.value
    return {"nodes": nodes[: c * 8].copy(), "node_count": c, "instance_indices": perm[:n].copy(),
            "blas_roots": roots[:n].copy(), "transforms": tf[: n * 16].copy()}


def _material(m) -> Material:
    em = m.get("emissive", [0.0, 0.0, 0.0])
    return Material((C.c_float * 3)(*m.get("baseColor", [0.8, 0.8, 0.8])),
                    float(m.get("roughness", 1.0)), float(m.get("metallic", 0.0)),
                    (C.c_float * 3)(*em), float(m.get("emissiveIntensity", 1.0 if any(em) else 0.0)),
                    1 if m.get("doubleSided", False) else 0,
                    float(m.get("clearcoat", 0.0)), float(m.get("clearcoatRoughness", 0.0)),
                    float(m.get("sheen", 0.0)), float(m.get("sheenRoughness", 0.0)),
                    (C.c_float * 3)(*m.get("sheenColor", [0.0, 0.0, 0.0])),
                    float(m.get("anisotropy", 0.0)), float(m.get("anisotropyRotation", 0.0)), 1,
                    (C.c_float * 3)(*m.get("specularColor", [1.0, 1.0, 1.0])),
                    float(m.get("specularIntensity", 1.0)))


class Engine:
    """Thin owner of a vt_engine (one per GPU device)."""

    def __init__(self, device: int = 0, max_bounces: int = 8, sampling: int = 0, light_selection: int = 0,
                 bvh_traversal: int = 0, build_threads: int = 16):
        """sampling: 0 PCG, 1 Owen-scrambled Sobol; light_selection: 0 light
        tree (two or more lights), 1 uniform; bvh_traversal: 0 binary BVH,
        1 cwbvh-closest (8-wide compressed BVH for closest hits);
        build_threads: host threads of the scene builder (0: all)."""
        self._h = C.c_void_p()
        opts = Options(max_bounces, 0, sampling, light_selection, bvh_traversal)
        check(lib().vt_create(C.c_int32(device), C.byref(opts), C.byref(self._h)))
        check(lib().vt_set_build_threads(self._h, C.c_uint32(build_threads)), self._h)

    def inject_failure(self, flags):
        """Test hook: make the next vt_set_scene fail at the chosen step (bit
        1 packing throws std::bad_alloc, 2 the shading-data upload fails, 4
        its preparation throws)."""
        check(lib().vt_debug_inject_failure(self._h, C.c_uint32(flags)), self._h)

    def close(self):
        if self._h:
            lib().vt_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def set_scene(self, scene):
        desc, _keep = _scene_desc(scene)
        check(lib().vt_set_scene(self._h, C.byref(desc)), self._h)

    def trace_closest(self, rays):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.zeros(rays.size, HIT_DTYPE)
        check(lib().vt_trace_closest(self._h, rays.ctypes.data, rays.size, hits.ctypes.data), self._h)
        return hits

    def trace_closest_counts(self, rays):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.zeros(rays.size, HIT_DTYPE)
        counts = np.zeros((rays.size, 4), np.uint32)
        check(lib().vt_trace_closest_counts(self._h, rays.ctypes.data, rays.size, hits.ctypes.data,
                                            counts.ctypes.data), self._h)
        return hits, counts

    def trace_any(self, rays):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        out = np.zeros(rays.size, np.uint32)
        check(lib().vt_trace_any(self._h, rays.ctypes.data, rays.size, out.ctypes.data), self._h)
        return out

    def set_node_width(self, width):
        """4: 4-wide trees collapsed from the binary BVH (default); 2: binary."""
        check(lib().vt_set_node_width(self._h, C.c_uint32(width)), self._h)

    def set_hit_filter(self, sided):
        """0: raw traversal (default); 1: the path tracer's sided wrappers."""
        check(lib().vt_set_hit_filter(self._h, C.c_uint32(sided)), self._h)

    def set_bvh_traversal(self, mode):
        """0: binary BVH; 1: cwbvh-closest (closest-hit rays only)."""
        check(lib().vt_set_bvh_traversal(self._h, C.c_uint32(mode)), self._h)

    def refit_primitive(self, primitive, positions, normals=None):
        """Moves primitive `primitive`'s vertices (same count) with its BVH
        topology kept: the device refit of its BLAS and the TLAS over its
        instances (vt_refit_primitive)."""
        pos = np.ascontiguousarray(positions, np.float32).ravel()
        nrm = None if normals is None else np.ascontiguousarray(normals, np.float32).ravel()
        check(lib().vt_refit_primitive(self._h, C.c_uint32(primitive), _p(pos, F32P), C.c_uint32(pos.size),
                                       _p(nrm, F32P), C.c_uint32(0 if nrm is None else nrm.size)), self._h)

    def refit_bvh_bounds_device(self, d_nodes, node_count, d_indices, index_words, index_stride, d_positions,
                                position_floats, position_stride):
        check(lib().vt_refit_bvh_bounds_device(self._h, C.c_void_p(d_nodes), C.c_uint32(node_count),
                                               C.c_void_p(d_indices), C.c_uint32(index_words),
                                               C.c_uint32(index_stride), C.c_void_p(d_positions),
                                               C.c_uint32(position_floats), C.c_uint32(position_stride)),
              self._h)

    def scene_field(self, field):
        """A field of the bound scene: a pack field name (host copy) or one of
        SCENE_BLAS_NODE64 / SCENE_BLAS_NODE128 / SCENE_TRIANGLES (device)."""
        fid = PACK_FIELDS.index(field) if isinstance(field, str) else field
        words = C.c_uint64(0)
        check(lib().vt_scene_copy_field(self._h, C.c_int32(fid), None, C.c_uint64(0), C.byref(words)), self._h)
        dtype = np.float32 if (field in FLOAT_PACK_FIELDS or fid == SCENE_TRIANGLES) else np.uint32
        out = np.zeros(max(words.value, 1), dtype)
        check(lib().vt_scene_copy_field(self._h, C.c_int32(fid), out.ctypes.data_as(C.c_void_p),
                                        C.c_uint64(out.size), C.byref(words)), self._h)
        return out[: words.value]

    def set_device_build(self, min_triangles):
        """Build the BLAS of meshes with at least `min_triangles` triangles on
        the GPU when a scene is bound (0: every BLAS on the host)."""
        check(lib().vt_set_device_build(self._h, C.c_uint32(min_triangles)), self._h)

    def set_shade_sort(self, enabled):
        check(lib().vt_set_shade_sort(self._h, C.c_uint32(1 if enabled else 0)), self._h)

    def set_trace_tuning(self, refill_min, steps, near_first=2, defer_min=12):
        check(lib().vt_set_trace_tuning(self._h, C.c_uint32(refill_min), C.c_uint32(steps),
                                        C.c_uint32(near_first), C.c_uint32(defer_min)), self._h)

    def selftest_fp32(self, a, b):
        a = np.ascontiguousarray(a, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        q = np.zeros_like(a)
        r = np.zeros_like(a)
        check(lib().vt_selftest_fp32(self._h, _p(a, F32P), _p(b, F32P), _p(q, F32P), _p(r, F32P),
                                     C.c_uint32(a.size)), self._h)
        return q, r

    # frames
    def render_frame(self, inv_view_proj, width, height, frame_index, frame_seed=0x9E3779B9,
                     max_bounces=0, samples_target=0):
        f = Frame((C.c_float * 16)(*np.asarray(inv_view_proj, np.float32).ravel()), width, height,
                  frame_index, frame_seed, max_bounces, samples_target)
        check(lib().vt_render_frame(self._h, C.byref(f)), self._h)

    def capture(self, width, height):
        out = np.zeros((height, width, 4), np.float32)
        check(lib().vt_capture(self._h, out.ctypes.data_as(F32P), width, height), self._h)
        return out

    def capture_output(self, width, height, tonemap=0, exposure=1.0, color_space=0):
        """captureFrame({colorSpace: 'output'}): tonemapped, OETF-encoded
        running mean rounded to binary16 (tonemap 0 aces .. 4 none;
        color_space 0 srgb, 1 linear)."""
        out = np.zeros((height, width, 4), np.float32)
        pr = Present(tonemap, exposure, color_space)
        check(lib().vt_capture_output(self._h, out.ctypes.data_as(F32P), C.c_uint32(width), C.c_uint32(height),
                                      C.byref(pr)), self._h)
        return out

    def capture_aux(self, which, width, height):
        """First-hit normal-depth (0) or albedo (1) of the latest frame, or
        the accumulated luminance variance (2); 4 floats per pixel."""
        out = np.zeros((height, width, 4), np.float32)
        check(lib().vt_capture_aux(self._h, C.c_uint32(which), out.ctypes.data_as(F32P), C.c_uint32(width),
                                   C.c_uint32(height)), self._h)
        return out

    def capture_signatures(self, width, height):
        out = np.zeros((height, width), np.uint64)
        check(lib().vt_capture_signatures(self._h, out.ctypes.data_as(C.POINTER(C.c_uint64)),
                                          width, height), self._h)
        return out

    def reset(self):
        check(lib().vt_reset(self._h), self._h)

    def frame_time_ms(self):
        ms = C.c_float()
        check(lib().vt_frame_time(self._h, C.byref(ms)), self._h)
        return ms.value

    # ---- neural radiance cache ----
    def nrc_param_count(self):
        n = C.c_uint32()
        check(lib().vt_nrc_param_count(C.byref(n)))
        return n.value

    def nrc_reset(self, learning_rate=1e-2, beta1=0.9, beta2=0.999, epsilon=1e-8, seed=1,
                  table_learning_rate=1e-1):
        o = NrcOptions(learning_rate, beta1, beta2, epsilon, seed, table_learning_rate)
        check(lib().vt_nrc_reset(self._h, C.byref(o)), self._h)

    def nrc_set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        check(lib().vt_nrc_set_weights(self._h, w.ctypes.data_as(F32P), C.c_uint32(w.size)), self._h)

    def nrc_get_weights(self):
        w = np.zeros(self.nrc_param_count(), np.float32)
        check(lib().vt_nrc_get_weights(self._h, w.ctypes.data_as(F32P), C.c_uint32(w.size)), self._h)
        return w

    def nrc_infer(self, features):
        f = np.ascontiguousarray(features, np.float32).reshape(-1, NRC_FEATURES)
        out = np.zeros((f.shape[0], 3), np.float32)
        check(lib().vt_nrc_infer(self._h, f.ctypes.data_as(F32P), C.c_uint32(f.shape[0]),
                                 out.ctypes.data_as(F32P)), self._h)
        return out

    def nrc_train_step(self, features, targets):
        f = np.ascontiguousarray(features, np.float32).reshape(-1, NRC_FEATURES)
        t = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
        loss = C.c_float()
        check(lib().vt_nrc_train_step(self._h, f.ctypes.data_as(F32P), t.ctypes.data_as(F32P),
                                      C.c_uint32(f.shape[0]), C.byref(loss)), self._h)
        return loss.value

    def nrc_gradient(self, features, targets):
        f = np.ascontiguousarray(features, np.float32).reshape(-1, NRC_FEATURES)
        t = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
        g = np.zeros(self.nrc_param_count(), np.float32)
        loss = C.c_float()
        check(lib().vt_nrc_gradient(self._h, f.ctypes.data_as(F32P), t.ctypes.data_as(F32P),
                                    C.c_uint32(f.shape[0]), g.ctypes.data_as(F32P), C.byref(loss)), self._h)
        return loss.value, g

    def nrc_infer_device(self, d_features, n, d_rgb):
        check(lib().vt_nrc_infer_device(self._h, C.c_void_p(d_features), C.c_uint32(n), C.c_void_p(d_rgb)), self._h)

    def nrc_train_step_device(self, d_features, d_targets, n):
        check(lib().vt_nrc_train_step_device(self._h, C.c_void_p(d_features), C.c_void_p(d_targets),
                                             C.c_uint32(n)), self._h)

    def nrc_render_config(self, enabled, train_paths=65536):
        check(lib().vt_nrc_render_config(self._h, C.c_int32(1 if enabled else 0), C.c_uint32(train_paths)), self._h)

    def nrc_teacher_vertices(self, vertices=4):
        """Shaded vertices of a self-training path before its terminal cache
        query (0: full-length training paths)."""
        check(lib().vt_nrc_teacher_vertices(self._h, C.c_uint32(vertices)), self._h)

    def nrc_timing(self):
        a, b = C.c_float(), C.c_float()
        check(lib().vt_nrc_timing(self._h, C.byref(a), C.byref(b)), self._h)
        return a.value, b.value

    def nrc_last_loss(self):
        loss = C.c_float()
        check(lib().vt_nrc_last_loss(self._h, C.byref(loss)), self._h)
        return loss.value

    def render_info(self):
        n = C.c_uint32()
        bias = C.c_float()
        check(lib().vt_render_info(self._h, C.byref(n), C.byref(bias)), self._h)
        return {"emitter_count": n.value, "ray_bias": bias.value}

    def light_tree(self):
        """The bound scene's packed NEE light tree, (nodes, 16) float32."""
        n = C.c_uint32()
        check(lib().vt_light_tree(self._h, None, 0, C.byref(n)), self._h)
        out = np.zeros((n.value, 16), np.float32)
        if n.value:
            check(lib().vt_light_tree(self._h, out.ctypes.data_as(C.POINTER(C.c_float)), n, C.byref(n)), self._h)
        return out

    # device-resident plumbing
    def alloc(self, nbytes):
        ptr = C.c_void_p()
        check(lib().vt_device_alloc(self._h, C.c_size_t(nbytes), C.byref(ptr)), self._h)
        return ptr.value

    def free(self, ptr):
        check(lib().vt_device_free(self._h, C.c_void_p(ptr)), self._h)

    def h2d(self, dptr, arr):
        arr = np.ascontiguousarray(arr)
        check(lib().vt_memcpy_h2d(self._h, C.c_void_p(dptr), arr.ctypes.data, arr.nbytes), self._h)

    def d2d(self, dst, src, nbytes):
        check(lib().vt_memcpy_d2d(self._h, C.c_void_p(dst), C.c_void_p(src), C.c_size_t(nbytes)), self._h)

    def d2h(self, arr, dptr):
        check(lib().vt_memcpy_d2h(self._h, arr.ctypes.data, C.c_void_p(dptr), arr.nbytes), self._h)

    def trace_closest_device(self, d_rays, n, d_hits):
        check(lib().vt_trace_closest_device(self._h, C.c_void_p(d_rays), n, C.c_void_p(d_hits)),
              self._h)

    def trace_closest_device_noalloc(self, d_rays, n, d_hits):
        check(lib().vt_trace_closest_device_noalloc(self._h, C.c_void_p(d_rays), C.c_uint32(n),
                                                    C.c_void_p(d_hits)), self._h)

    def trace_any_device(self, d_rays, n, d_out):
        check(lib().vt_trace_any_device(self._h, C.c_void_p(d_rays), n, C.c_void_p(d_out)), self._h)

    def synchronize(self):
        check(lib().vt_synchronize(self._h), self._h)

    def stats(self):
        s = FrameStats()
        check(lib().vt_stats(self._h, C.byref(s)), self._h)
        return {k: getattr(s, k) for k, _ in FrameStats._fields_}


def _scene_desc(scene):
    """(SceneDesc, keep-alive objects) of a scene dict."""
    prims = []
    mats = []
    for i, p in enumerate(scene["primitives"]):
        prims.append(dict(p, material_id=i))
        mats.append(_material(p.get("material", {})))
    arr, keep = _prim_structs(prims)
    marr = (Material * max(1, len(mats)))(*mats)
    env, tex = _environment(scene)
    ems = scene.get("emitters", [])
    earr = (Emitter * max(1, len(ems)))(*[_emitter(e) for e in ems])
    desc = SceneDesc(arr, len(prims), marr, len(mats), env, float(scene.get("scene_radius", 0.0)), earr, len(ems))
    return desc, (arr, keep, marr, earr, tex)


class Batch:
    """Several scenes traced in one persistent launch (vt_batch_*)."""

    def __init__(self, engine):
        self._eng = engine
        self._h = C.c_void_p()
        check(lib().vt_batch_create(engine.handle, C.byref(self._h)), engine.handle)

    def close(self):
        if self._h:
            lib().vt_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_scene(self, scene):
        desc, _keep = _scene_desc(scene)
        idx = C.c_uint32()
        check(lib().vt_batch_add_scene(self._h, C.byref(desc), C.byref(idx)), self._eng.handle)
        return idx.value

    @staticmethod
    def _offsets(offsets):
        o = np.ascontiguousarray(offsets, np.uint32)
        return o, o.ctypes.data_as(U32P)

    def trace_closest_device(self, d_rays, offsets, d_hits, timed=True):
        o, op = self._offsets(offsets)
        ms = C.c_float()
        check(lib().vt_batch_trace_closest_device(self._h, C.c_void_p(d_rays), op, C.c_void_p(d_hits),
                                                  C.byref(ms) if timed else None), self._eng.handle)
        return ms.value

    def capture(self, d_rays, offsets, d_hits):
        o, op = self._offsets(offsets)
        check(lib().vt_batch_capture(self._h, C.c_void_p(d_rays), op, C.c_void_p(d_hits)), self._eng.handle)

    def replay(self):
        ms = C.c_float()
        check(lib().vt_batch_replay(self._h, C.byref(ms)), self._eng.handle)
        return ms.value

    def set_tuning(self, refill_min, steps, near_first, defer_min, blocks_per_cu):
        t = Tuning(refill_min, steps, near_first, defer_min, blocks_per_cu, 0.0)
        check(lib().vt_batch_set_tuning(self._h, C.byref(t)), self._eng.handle)

    def autotune(self, d_rays, offsets, d_hits, repeats=3):
        o, op = self._offsets(offsets)
        t = Tuning()
        check(lib().vt_batch_autotune(self._h, C.c_void_p(d_rays), op, C.c_void_p(d_hits), C.c_uint32(repeats),
                                      C.byref(t)), self._eng.handle)
        return {"refill_min": t.refill_min, "steps": t.steps, "near_first": t.near_first,
                "defer_min": t.defer_min, "blocks_per_cu": t.blocks_per_cu, "ms": t.ms}


class RenderBatch:
    """Several engines' frames rendered together (vt_render_batch_*): every
    bounce's rays of all members traced in one persistent launch, on the
    host engine's stream; capture() records a frame into a HIP graph that
    replay() launches for later frame indices."""

    def __init__(self, host):
        self._host = host
        self._h = C.c_void_p()
        self._members = []
        check(lib().vt_render_batch_create(host.handle, C.byref(self._h)), host.handle)

    def close(self):
        if self._h:
            lib().vt_render_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, engine):
        check(lib().vt_render_batch_add(self._h, engine.handle), self._host.handle)
        self._members.append(engine)

    @staticmethod
    def _frames(frames):
        arr = (Frame * len(frames))()
        for k, fr in enumerate(frames):
            arr[k] = Frame((C.c_float * 16)(*np.asarray(fr["inv_view_proj"], np.float32).ravel()), fr["width"],
                           fr["height"], fr["frame_index"], fr.get("frame_seed", 0x9E3779B9),
                           fr.get("max_bounces", 0), fr.get("samples_target", 0))
        return arr

    def frame(self, frames):
        """frames: one dict per member (inv_view_proj, width, height,
        frame_index, frame_seed, max_bounces)."""
        arr = self._frames(frames)
        check(lib().vt_render_batch_frame(self._h, C.cast(arr, C.c_void_p)), self._host.handle)

    def capture(self, frames):
        arr = self._frames(frames)
        check(lib().vt_render_batch_capture(self._h, C.cast(arr, C.c_void_p)), self._host.handle)

    def replay(self, frame_index):
        check(lib().vt_render_batch_replay(self._h, C.c_uint32(frame_index)), self._host.handle)
