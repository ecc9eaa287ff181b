// addon.cpp -- Node N-API binding of libvthip.so.
//
// The JavaScript backend (vitrum_amd/js/index.js) implements the reference's
// Engine contract (packages/core/src/engine/index.ts:109-486) on top of these
// functions; each one forwards to one entry point of include/vitrum_hip.h.
// Typed-array backing stores are borrowed for the duration of a call only.
// Library status codes become the JavaScript error classes the reference
// throws: VT_ERR_INVALID_ARGUMENT -> TypeError, VT_ERR_RANGE -> RangeError,
// anything else -> Error with `code` = "VT_ERR_*".
#include <node_api.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/vitrum_hip.h"

namespace {

struct Handle {
  vt_engine* eng = nullptr;
};

const char* status_code(vt_status s) {
  switch (s) {
    case VT_ERR_INVALID_ARGUMENT: return "VT_ERR_INVALID_ARGUMENT";
    case VT_ERR_RANGE: return "VT_ERR_RANGE";
    case VT_ERR_BUILD: return "VT_ERR_BUILD";
    case VT_ERR_HIP: return "VT_ERR_HIP";
    case VT_ERR_NO_DEVICE: return "VT_ERR_NO_DEVICE";
    case VT_ERR_STATE: return "VT_ERR_STATE";
    case VT_ERR_OUT_OF_MEMORY: return "VT_ERR_OUT_OF_MEMORY";
    default: return "VT_ERR_UNKNOWN";
  }
}

// Throws the JavaScript error for a failed call; returns true when it threw.
bool throw_status(napi_env env, const vt_engine* eng, vt_status s) {
  if (s == VT_OK) return false;
  char msg[1024] = {0};
  vt_last_error(eng, msg, sizeof msg);
  if (!msg[0]) std::snprintf(msg, sizeof msg, "vitrum-hip call failed (%s)", status_code(s));
  if (s == VT_ERR_INVALID_ARGUMENT)
    napi_throw_type_error(env, status_code(s), msg);
  else if (s == VT_ERR_RANGE)
    napi_throw_range_error(env, status_code(s), msg);
  else
    napi_throw_error(env, status_code(s), msg);
  return true;
}

bool throw_type(napi_env env, const char* msg) {
  napi_throw_type_error(env, "VT_ERR_INVALID_ARGUMENT", msg);
  return true;
}

#define NAPI_OK(call)                                              \
  do {                                                             \
    if ((call) != napi_ok) {                                       \
      napi_throw_error(env, nullptr, "N-API call failed: " #call); \
      return nullptr;                                              \
    }                                                              \
  } while (0)

std::vector<napi_value> args_of(napi_env env, napi_callback_info info, size_t n) {
  std::vector<napi_value> v(n, nullptr);
  size_t argc = n;
  napi_get_cb_info(env, info, &argc, v.data(), nullptr, nullptr);
  for (size_t i = argc; i < n; ++i) napi_get_undefined(env, &v[i]);
  return v;
}

bool is_nullish(napi_env env, napi_value v) {
  napi_valuetype t;
  if (napi_typeof(env, v, &t) != napi_ok) return true;
  return t == napi_undefined || t == napi_null;
}

bool get_u32(napi_env env, napi_value obj, const char* key, uint32_t def, uint32_t* out) {
  napi_value v;
  bool has = false;
  napi_has_named_property(env, obj, key, &has);
  if (!has) {
    *out = def;
    return true;
  }
  napi_get_named_property(env, obj, key, &v);
  if (is_nullish(env, v)) {
    *out = def;
    return true;
  }
  double d;
  if (napi_get_value_double(env, v, &d) != napi_ok || !(d >= 0) || d > 4294967295.0 || d != (double)(uint32_t)d) {
    std::string m = std::string("'") + key + "' must be an unsigned 32-bit integer";
    throw_type(env, m.c_str());
    return false;
  }
  *out = (uint32_t)d;
  return true;
}

bool get_f32(napi_env env, napi_value obj, const char* key, float def, float* out) {
  napi_value v;
  bool has = false;
  napi_has_named_property(env, obj, key, &has);
  if (!has) {
    *out = def;
    return true;
  }
  napi_get_named_property(env, obj, key, &v);
  if (is_nullish(env, v)) {
    *out = def;
    return true;
  }
  double d;
  if (napi_get_value_double(env, v, &d) != napi_ok) {
    std::string m = std::string("'") + key + "' must be a number";
    throw_type(env, m.c_str());
    return false;
  }
  *out = (float)d;
  return true;
}

// Reads `n` numbers from a JS array or typed array property into out[].
bool get_vec(napi_env env, napi_value obj, const char* key, uint32_t n, const float* def, float* out) {
  napi_value v;
  bool has = false;
  napi_has_named_property(env, obj, key, &has);
  if (has) napi_get_named_property(env, obj, key, &v);
  if (!has || is_nullish(env, v)) {
    for (uint32_t i = 0; i < n; ++i) out[i] = def[i];
    return true;
  }
  for (uint32_t i = 0; i < n; ++i) {
    napi_value e;
    double d;
    if (napi_get_element(env, v, i, &e) != napi_ok || napi_get_value_double(env, e, &d) != napi_ok) {
      std::string m = std::string("'") + key + "' must hold " + std::to_string(n) + " numbers";
      throw_type(env, m.c_str());
      return false;
    }
    out[i] = (float)d;
  }
  return true;
}

// Binary64 number / n-vector properties (JS numbers kept whole).
bool get_f64(napi_env env, napi_value obj, const char* key, double def, double* out) {
  napi_value v;
  bool has = false;
  napi_has_named_property(env, obj, key, &has);
  if (has) napi_get_named_property(env, obj, key, &v);
  if (!has || is_nullish(env, v)) {
    *out = def;
    return true;
  }
  if (napi_get_value_double(env, v, out) != napi_ok) {
    std::string m = std::string("'") + key + "' must be a number";
    throw_type(env, m.c_str());
    return false;
  }
  return true;
}

bool get_dvec(napi_env env, napi_value obj, const char* key, uint32_t n, const double* def, double* out) {
  napi_value v;
  bool has = false;
  napi_has_named_property(env, obj, key, &has);
  if (has) napi_get_named_property(env, obj, key, &v);
  if (!has || is_nullish(env, v)) {
    for (uint32_t i = 0; i < n; ++i) out[i] = def[i];
    return true;
  }
  for (uint32_t i = 0; i < n; ++i) {
    napi_value e;
    if (napi_get_element(env, v, i, &e) != napi_ok || napi_get_value_double(env, e, &out[i]) != napi_ok) {
      std::string m = std::string("'") + key + "' must hold " + std::to_string(n) + " numbers";
      throw_type(env, m.c_str());
      return false;
    }
  }
  return true;
}

// Borrows the backing store of a typed-array property of the given type.
bool get_typed(napi_env env, napi_value obj, const char* key, napi_typedarray_type want, bool optional,
               void** data, size_t* length) {
  napi_value v;
  bool has = false;
  napi_has_named_property(env, obj, key, &has);
  if (has) napi_get_named_property(env, obj, key, &v);
  if (!has || is_nullish(env, v)) {
    *data = nullptr;
    *length = 0;
    if (optional) return true;
    std::string m = std::string("'") + key + "' is required";
    throw_type(env, m.c_str());
    return false;
  }
  bool is_ta = false;
  napi_is_typedarray(env, v, &is_ta);
  napi_typedarray_type t;
  napi_value ab;
  size_t off;
  if (!is_ta || napi_get_typedarray_info(env, v, &t, length, data, &ab, &off) != napi_ok || t != want) {
    std::string m = std::string("'") + key + "' has the wrong typed-array type";
    throw_type(env, m.c_str());
    return false;
  }
  return true;
}

Handle* handle_of(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    throw_type(env, "expected an engine handle");
    return nullptr;
  }
  Handle* h = static_cast<Handle*>(p);
  if (!h->eng) {
    napi_throw_error(env, "VT_ERR_STATE", "engine has been disposed");
    return nullptr;
  }
  return h;
}

void finalize_handle(napi_env, void* data, void*) {
  Handle* h = static_cast<Handle*>(data);
  if (h->eng) vt_destroy(h->eng);
  delete h;
}

// create(deviceOrdinal, {maxBounces, maxSamplesPerPixel, sampling}) -> handle
napi_value Create(napi_env env, napi_callback_info info) {
  auto a = args_of(env, info, 2);
  int32_t dev = 0;
  if (!is_nullish(env, a[0]) && napi_get_value_int32(env, a[0], &dev) != napi_ok) {
    throw_type(env, "device ordinal must be an integer");
    return nullptr;
  }
  vt_options o{};
  if (!is_nullish(env, a[1])) {
    if (!get_u32(env, a[1], "maxBounces", 0, &o.max_bounces)) return nullptr;
    if (!get_u32(env, a[1], "maxSamplesPerPixel", 0, &o.max_samples_per_pixel)) return nullptr;
    if (!get_u32(env, a[1], "sampling", 0, &o.sampling)) return nullptr;
    if (!get_u32(env, a[1], "lightSelection", 0, &o.light_selection)) return nullptr;
    if (!get_u32(env, a[1], "bvhTraversal", 0, &o.bvh_traversal)) return nullptr;
  }
  vt_engine* eng = nullptr;
  if (throw_status(env, nullptr, vt_create(dev, &o, &eng))) return nullptr;
  Handle* h = new Handle{eng};
  napi_value out;
  NAPI_OK(napi_create_external(env, h, finalize_handle, nullptr, &out));
  return out;
}

napi_value Destroy(napi_env env, napi_callback_info info) {
  auto a = args_of(env, info, 1);
  void* p = nullptr;
  if (napi_get_value_external(env, a[0], &p) == napi_ok && p) {
    Handle* h = static_cast<Handle*>(p);
    if (h->eng) vt_destroy(h->eng);
    h->eng = nullptr;
  }
  return nullptr;
}

// setScene(handle, {primitives, materials, environment, sceneRadius})
napi_value SetScene(napi_env env, napi_callback_info info) {
  auto a = args_of(env, info, 2);
  Handle* h = handle_of(env, a[0]);
  if (!h) return nullptr;
  napi_value desc = a[1], prims, mats;
  bool is_arr = false;
  if (is_nullish(env, desc)) {
    throw_type(env, "scene description is required");
    return nullptr;
  }
  NAPI_OK(napi_get_named_property(env, desc, "primitives", &prims));
  napi_is_array(env, prims, &is_arr);
  if (!is_arr) {
    throw_type(env, "'primitives' must be an array");
    return nullptr;
  }
  uint32_t np = 0, nm = 0;
  napi_get_array_length(env, prims, &np);
  std::vector<vt_primitive> vp(np);
  for (uint32_t i = 0; i < np; ++i) {
    napi_value pr;
    NAPI_OK(napi_get_element(env, prims, i, &pr));
    vt_primitive& q = vp[i];
    std::memset(&q, 0, sizeof q);
    if (!get_u32(env, pr, "kind", VT_PRIM_MESH, &q.kind)) return nullptr;
    if (!get_u32(env, pr, "material", i, &q.material)) return nullptr;
    void* d;
    size_t n;
    if (!get_typed(env, pr, "positions", napi_float32_array, false, &d, &n)) return nullptr;
    q.positions = static_cast<const float*>(d);
    q.position_floats = (uint32_t)n;
    if (!get_typed(env, pr, "normals", napi_float32_array, true, &d, &n)) return nullptr;
    q.normals = static_cast<const float*>(d);
    q.normal_floats = (uint32_t)n;
    if (!get_typed(env, pr, "indices", napi_uint32_array, true, &d, &n)) return nullptr;
    q.indices = static_cast<const uint32_t*>(d);
    q.index_count = (uint32_t)n;
    if (!get_typed(env, pr, "transforms", napi_float32_array, true, &d, &n)) return nullptr;
    q.transforms = static_cast<const float*>(d);
    q.transform_count = (uint32_t)(n / 16);
    if (n % 16) {
      throw_type(env, "'transforms' must hold whole column-major mat4s");
      return nullptr;
    }
  }
  NAPI_OK(napi_get_named_property(env, desc, "materials", &mats));
  napi_is_array(env, mats, &is_arr);
  if (!is_arr) {
    throw_type(env, "'materials' must be an array");
    return nullptr;
  }
  napi_get_array_length(env, mats, &nm);
  std::vector<vt_material> vm(nm);
  const float white[3] = {0.8f, 0.8f, 0.8f}, zero[3] = {0, 0, 0};
  for (uint32_t i = 0; i < nm; ++i) {
    napi_value m;
    NAPI_OK(napi_get_element(env, mats, i, &m));
    vt_material& q = vm[i];
    std::memset(&q, 0, sizeof q);
    if (!get_vec(env, m, "baseColor", 3, white, q.base_color)) return nullptr;
    if (!get_f32(env, m, "roughness", 0.5f, &q.roughness)) return nullptr;
    if (!get_f32(env, m, "metallic", 0.0f, &q.metallic)) return nullptr;
    if (!get_vec(env, m, "emissive", 3, zero, q.emissive)) return nullptr;
    if (!get_f32(env, m, "emissiveIntensity", 1.0f, &q.emissive_intensity)) return nullptr;
    bool ds = false;
    napi_value v;
    bool has = false;
    napi_has_named_property(env, m, "doubleSided", &has);
    if (has && napi_get_named_property(env, m, "doubleSided", &v) == napi_ok) napi_get_value_bool(env, v, &ds);
    q.double_sided = ds ? 1u : 0u;
    // extension lobes (MaterialSpec clearcoat / sheen / anisotropy / specular)
    if (!get_f32(env, m, "clearcoat", 0.0f, &q.clearcoat)) return nullptr;
    if (!get_f32(env, m, "clearcoatRoughness", 0.0f, &q.clearcoat_roughness)) return nullptr;
    if (!get_f32(env, m, "sheen", 0.0f, &q.sheen)) return nullptr;
    if (!get_f32(env, m, "sheenRoughness", 0.0f, &q.sheen_roughness)) return nullptr;
    if (!get_vec(env, m, "sheenColor", 3, zero, q.sheen_color)) return nullptr;
    if (!get_f32(env, m, "anisotropy", 0.0f, &q.anisotropy)) return nullptr;
    if (!get_f32(env, m, "anisotropyRotation", 0.0f, &q.anisotropy_rotation)) return nullptr;
    const float one[3] = {1, 1, 1};
    q.has_specular = 1u;
    if (!get_vec(env, m, "specularColor", 3, one, q.specular_color)) return nullptr;
    if (!get_f32(env, m, "specularIntensity", 1.0f, &q.specular_intensity)) return nullptr;
  }
  vt_scene_desc sd{};
  sd.primitives = vp.data();
  sd.primitive_count = np;
  sd.materials = vm.data();
  sd.material_count = nm;
  napi_value envv;
  bool has_env = false;
  napi_has_named_property(env, desc, "environment", &has_env);
  if (has_env && napi_get_named_property(env, desc, "environment", &envv) == napi_ok && !is_nullish(env, envv)) {
    // {kind: 1, radiance} or {kind: 2, texels: Float32Array, width, height,
    //  channels, intensity, rotationY}; the engine copies the texels.
    if (!get_u32(env, envv, "kind", 0, &sd.environment.kind)) return nullptr;
    if (!get_vec(env, envv, "radiance", 3, zero, sd.environment.radiance)) return nullptr;
    void* td = nullptr;
    size_t tn = 0;
    if (!get_typed(env, envv, "texels", napi_float32_array, true, &td, &tn)) return nullptr;
    if (!get_u32(env, envv, "width", 0, &sd.environment.width)) return nullptr;
    if (!get_u32(env, envv, "height", 0, &sd.environment.height)) return nullptr;
    if (!get_u32(env, envv, "channels", 0, &sd.environment.channels)) return nullptr;
    if (!get_f32(env, envv, "intensity", 1.0f, &sd.environment.intensity)) return nullptr;
    if (!get_f32(env, envv, "rotationY", 0.0f, &sd.environment.rotation_y)) return nullptr;
    if (td && (uint64_t)sd.environment.width * sd.environment.height * sd.environment.channels != tn) {
      throw_type(env, "'texels' must hold width x height x channels floats");
      return nullptr;
    }
    sd.environment.texels = (const float*)td;
    // {kind: 3, sunDirection, turbidity, rayleigh, mieCoefficient,
    //  mieDirectionalG, intensity} (defaults of environmentPacking.ts:245-262)
    const double up[3] = {0.0, 1.0, 0.0};
    if (!get_dvec(env, envv, "sunDirection", 3, up, sd.environment.sun_direction)) return nullptr;
    if (!get_f64(env, envv, "turbidity", 2.0, &sd.environment.turbidity)) return nullptr;
    if (!get_f64(env, envv, "rayleigh", 1.0, &sd.environment.rayleigh)) return nullptr;
    if (!get_f64(env, envv, "mieCoefficient", 0.005, &sd.environment.mie_coefficient)) return nullptr;
    if (!get_f64(env, envv, "mieDirectionalG", 0.8, &sd.environment.mie_directional_g)) return nullptr;
  }
  if (!get_f32(env, desc, "sceneRadius", 0.0f, &sd.scene_radius)) return nullptr;
  // emitters: [{kind, color, intensity, castShadow, position, direction, normal,
  //             uAxis, vAxis, radius, angle, penumbra, distance, decay, angularDiameter}]
  std::vector<vt_emitter> ve;
  bool has_em = false;
  napi_has_named_property(env, desc, "emitters", &has_em);
  if (has_em) {
    napi_value ems;
    NAPI_OK(napi_get_named_property(env, desc, "emitters", &ems));
    bool em_arr = false;
    napi_is_array(env, ems, &em_arr);
    if (!em_arr) {
      throw_type(env, "'emitters' must be an array");
      return nullptr;
    }
    uint32_t ne = 0;
    napi_get_array_length(env, ems, &ne);
    ve.resize(ne);
    const float down[3] = {0, -1, 0};
    for (uint32_t i = 0; i < ne; ++i) {
      napi_value e;
      NAPI_OK(napi_get_element(env, ems, i, &e));
      vt_emitter& q = ve[i];
      std::memset(&q, 0, sizeof q);
      const float one[3] = {1, 1, 1};
      if (!get_u32(env, e, "kind", 0, &q.kind)) return nullptr;
      if (!get_vec(env, e, "color", 3, one, q.color)) return nullptr;
      if (!get_f32(env, e, "intensity", 1.0f, &q.intensity)) return nullptr;
      bool cs = true, has_cs = false;
      napi_has_named_property(env, e, "castShadow", &has_cs);
      if (has_cs) {
        napi_value v;
        if (napi_get_named_property(env, e, "castShadow", &v) == napi_ok) napi_get_value_bool(env, v, &cs);
      }
      q.cast_shadow = cs ? 1u : 0u;
      if (!get_vec(env, e, "position", 3, zero, q.position)) return nullptr;
      if (!get_vec(env, e, "direction", 3, down, q.direction)) return nullptr;
      if (!get_vec(env, e, "normal", 3, down, q.normal)) return nullptr;
      if (!get_vec(env, e, "uAxis", 3, zero, q.u_axis)) return nullptr;
      if (!get_vec(env, e, "vAxis", 3, zero, q.v_axis)) return nullptr;
      if (!get_f32(env, e, "radius", 0.0f, &q.radius)) return nullptr;
      if (!get_f32(env, e, "angle", 0.0f, &q.angle)) return nullptr;
      if (!get_f32(env, e, "penumbra", 0.0f, &q.penumbra)) return nullptr;
      if (!get_f32(env, e, "distance", 0.0f, &q.distance)) return nullptr;
      if (!get_f32(env, e, "decay", 2.0f, &q.decay)) return nullptr;
      if (!get_f32(env, e, "angularDiameter", 0.0f, &q.angular_diameter)) return nullptr;
    }
  }
  sd.emitters = ve.data();
  sd.emitter_count = (uint32_t)ve.size();
  if (throw_status(env, h->eng, vt_set_scene(h->eng, &sd))) return nullptr;
  return nullptr;
}

// renderFrame(handle, {invViewProj, width, height, frameIndex, frameSeed, maxBounces, samplesTarget})
napi_value RenderFrame(napi_env env, napi_callback_info info) {
  auto a = args_of(env, info, 2);
  Handle* h = handle_of(env, a[0]);
  if (!h) return nullptr;
  if (is_nullish(env, a[1])) {
    throw_type(env, "frame description is required");
    return nullptr;
  }
  vt_frame f{};
  void* d;
  size_t n;
  if (!get_typed(env, a[1], "invViewProj", napi_float32_array, false, &d, &n)) return nullptr;
  if (n != 16) {
    throw_type(env, "'invViewProj' must be a Float32Array(16)");
    return nullptr;
  }
  std::memcpy(f.inv_view_proj, d, 64);
  if (!get_u32(env, a[1], "width", 0, &f.width)) return nullptr;
  if (!get_u32(env, a[1], "height", 0, &f.height)) return nullptr;
  if (!get_u32(env, a[1], "frameIndex", 0, &f.frame_index)) return nullptr;
  if (!get_u32(env, a[1], "frameSeed", 0, &f.frame_seed)) return nullptr;
  if (!get_u32(env, a[1], "maxBounces", 0, &f.max_bounces)) return nullptr;
  if (!get_u32(env, a[1], "samplesTarget", 0, &f.samples_target)) return nullptr;
  if (throw_status(env, h->eng, vt_render_frame(h->eng, &f))) return nullptr;
  return nullptr;
}

// Reads one frame description (renderFrame's second argument).
bool frame_of(napi_env env, napi_value v, vt_frame* f) {
  if (is_nullish(env, v)) {
    throw_type(env, "frame description is required");
    return false;
  }
  *f = vt_frame{};
  void* d;
  size_t n;
  if (!get_typed(env, v, "invViewProj", napi_float32_array, false, &d, &n)) return false;
  if (n != 16) {
    throw_type(env, "'invViewProj' must be a Float32Array(16)");
    return false;
  }
  std::memcpy(f->inv_view_proj, d, 64);
  return get_u32(env, v, "width", 0, &f->width) && get_u32(env, v, "height", 0, &f->height) &&
         get_u32(env, v, "frameIndex", 0, &f->frame_index) && get_u32(env, v, "frameSeed", 0, &f->frame_seed) &&
         get_u32(env, v, "maxBounces", 0, &f->max_bounces) && get_u32(env, v, "samplesTarget", 0, &f->samples_target);
}

// renderFrames([handle...], [frame...]): one frame of every engine, batched
// (vt_render_batch_*: each bounce's rays of all engines in one traversal
// launch, on the first engine's stream).
napi_value RenderFrames(napi_env env, napi_callback_info info) {
  auto a = args_of(env, info, 2);
  bool ha = false, fa = false;
  napi_is_array(env, a[0], &ha);
  napi_is_array(env, a[1], &fa);
  uint32_t nh = 0, nf = 0;
  if (!ha || !fa || napi_get_array_length(env, a[0], &nh) != napi_ok || napi_get_array_length(env, a[1], &nf) != napi_ok ||
      nh == 0 || nh != nf) {
    throw_type(env, "renderFrames([handle...], [frame...]) takes equally long, non-empty arrays");
    return nullptr;
  }
  std::vector<vt_engine*> engs(nh);
  std::vector<vt_frame> frames(nh);
  for (uint32_t i = 0; i < nh; ++i) {
    napi_value hv, fv;
    NAPI_OK(napi_get_element(env, a[0], i, &hv));
    NAPI_OK(napi_get_element(env, a[1], i, &fv));
    Handle* h = handle_of(env, hv);
    if (!h) return nullptr;
    engs[i] = h->eng;
    if (!frame_of(env, fv, &frames[i])) return nullptr;
  }
  vt_render_batch* b = nullptr;
  if (throw_status(env, engs[0], vt_render_batch_create(engs[0], &b))) return nullptr;
  vt_status s = VT_OK;
  for (uint32_t i = 0; i < nh && !s; ++i) s = vt_render_batch_add(b, engs[i]);
  if (!s) s = vt_render_batch_frame(b, frames.data());
  if (s) {
    throw_status(env, engs[0], s);
    vt_render_batch_destroy(b);
    return nullptr;
  }
  vt_render_batch_destroy(b);
  return nullptr;
}

// capture(handle, width, height) -> Float32Array (row-major RGBA, top-left origin)
napi_value Capture(napi_env env, napi_callback_info info) {
  auto a = args_of(env, info, 3);
  Handle* h = handle_of(env, a[0]);
  if (!h) return nullptr;
  uint32_t w = 0, hh = 0;
  if (napi_get_value_uint32(env, a[1], &w) != napi_ok || napi_get_value_uint32(env, a[2], &hh) != napi_ok) {
    throw_type(env, "width and height must be unsigned integers");
    return nullptr;
  }
  const size_t floats = (size_t)w * hh * 4;
  void* data = nullptr;
  napi_value ab, out;
  NAPI_OK(napi_create_arraybuffer(env, floats * 4, &data, &ab));
  if (throw_status(env, h->eng, vt_capture(h->eng, static_cast<float*>(data), w, hh))) return nullptr;
  NAPI_OK(napi_create_typedarray(env, napi_float32_array, floats, ab, 0, &out));
  return out;
}

// captureOutput(handle, width, height, tonemap, exposure, colorSpace) ->
// Float32Array: the present pass (tonemap, OETF, binary16 store).
napi_value CaptureOutput(napi_env env, napi_callback_info info) {
  auto a = args_of(env, info, 6);
  Handle* h = handle_of(env, a[0]);
  if (!h) return nullptr;
  uint32_t w = 0, hh = 0;
  vt_present pr{};
  double ex = 1.0;
  if (napi_get_value_uint32(env, a[1], &w) != napi_ok || napi_get_value_uint32(env, a[2], &hh) != napi_ok ||
      napi_get_value_uint32(env, a[3], &pr.tonemap) != napi_ok || napi_get_value_double(env, a[4], &ex) != napi_ok ||
      napi_get_value_uint32(env, a[5], &pr.color_space) != napi_ok) {
    throw_type(env, "captureOutput(handle, width, height, tonemap, exposure, colorSpace)");
    return nullptr;
  }
  pr.exposure = (float)ex;
  const size_t floats = (size_t)w * hh * 4;
  void* data = nullptr;
  napi_value ab, out;
  NAPI_OK(napi_create_arraybuffer(env, floats * 4, &data, &ab));
  if (throw_status(env, h->eng, vt_capture_output(h->eng, static_cast<float*>(data), w, hh, &pr))) return nullptr;
  NAPI_OK(napi_create_typedarray(env, napi_float32_array, floats, ab, 0, &out));
  return out;
}

// captureAux(handle, which, width, height) -> Float32Array: 0 normal-depth,
// 1 albedo, 2 variance.
napi_value CaptureAux(napi_env env, napi_callback_info info) {
  auto a = args_of(env, info, 4);
  Handle* h = handle_of(env, a[0]);
  if (!h) return nullptr;
  uint32_t which = 0, w = 0, hh = 0;
  if (napi_get_value_uint32(env, a[1], &which) != napi_ok || napi_get_value_uint32(env, a[2], &w) != napi_ok ||
      napi_get_value_uint32(env, a[3], &hh) != napi_ok) {
    throw_type(env, "captureAux(handle, which, width, height)");
    return nullptr;
  }
  const size_t floats = (size_t)w * hh * 4;
  void* data = nullptr;
  napi_value ab, out;
  NAPI_OK(napi_create_arraybuffer(env, floats * 4, &data, &ab));
  if (throw_status(env, h->eng, vt_capture_aux(h->eng, which, static_cast<float*>(data), w, hh))) return nullptr;
  NAPI_OK(napi_create_typedarray(env, napi_float32_array, floats, ab, 0, &out));
  return out;
}

napi_value Reset(napi_env env, napi_callback_info info) {
  auto a = args_of(env, info, 1);
  Handle* h = handle_of(env, a[0]);
  if (!h) return nullptr;
  throw_status(env, h->eng, vt_reset(h->eng));
  return nullptr;
}

// refitPrimitive(handle, primitiveIndex, {positions, normals?}): moves one
// primitive's vertices with its BVH topology kept (vt_refit_primitive).
napi_value RefitPrimitive(napi_env env, napi_callback_info info) {
  auto a = args_of(env, info, 3);
  Handle* h = handle_of(env, a[0]);
  if (!h) return nullptr;
  uint32_t index = 0;
  if (napi_get_value_uint32(env, a[1], &index) != napi_ok) {
    throw_type(env, "primitive index must be a number");
    return nullptr;
  }
  void *pos = nullptr, *nrm = nullptr;
  size_t npos = 0, nnrm = 0;
  if (!get_typed(env, a[2], "positions", napi_float32_array, false, &pos, &npos)) return nullptr;
  if (!get_typed(env, a[2], "normals", napi_float32_array, true, &nrm, &nnrm)) return nullptr;
  throw_status(env, h->eng,
               vt_refit_primitive(h->eng, index, (const float*)pos, (uint32_t)npos, (const float*)nrm, (uint32_t)nnrm));
  return nullptr;
}

napi_value Stats(napi_env env, napi_callback_info info) {
  auto a = args_of(env, info, 1);
  Handle* h = handle_of(env, a[0]);
  if (!h) return nullptr;
  vt_frame_stats s{};
  if (throw_status(env, h->eng, vt_stats(h->eng, &s))) return nullptr;
  float ms = 0;
  if (throw_status(env, h->eng, vt_frame_time(h->eng, &ms))) return nullptr;
  napi_value o, v;
  NAPI_OK(napi_create_object(env, &o));
  napi_create_double(env, s.last_trace_ms, &v);
  napi_set_named_property(env, o, "lastTraceMs", v);
  napi_create_double(env, ms, &v);
  napi_set_named_property(env, o, "lastFrameMs", v);
  napi_create_uint32(env, s.samples_accumulated, &v);
  napi_set_named_property(env, o, "samplesAccumulated", v);
  napi_create_uint32(env, s.last_ray_count, &v);
  napi_set_named_property(env, o, "lastRayCount", v);
  napi_create_double(env, (double)s.device_bytes, &v);
  napi_set_named_property(env, o, "deviceBytes", v);
  return o;
}

// traceClosest(handle, rays: Float32Array(8 n)) -> ArrayBuffer of n 20-byte hits
napi_value TraceClosest(napi_env env, napi_callback_info info) {
  auto a = args_of(env, info, 2);
  Handle* h = handle_of(env, a[0]);
  if (!h) return nullptr;
  bool is_ta = false;
  napi_is_typedarray(env, a[1], &is_ta);
  napi_typedarray_type t;
  size_t len = 0, off = 0;
  void* d = nullptr;
  napi_value ab0;
  if (!is_ta || napi_get_typedarray_info(env, a[1], &t, &len, &d, &ab0, &off) != napi_ok ||
      t != napi_float32_array || len % 8) {
    throw_type(env, "rays must be a Float32Array of 8 floats per ray");
    return nullptr;
  }
  const uint32_t n = (uint32_t)(len / 8);
  void* out = nullptr;
  napi_value ab;
  NAPI_OK(napi_create_arraybuffer(env, (size_t)n * sizeof(vt_hit), &out, &ab));
  if (throw_status(env, h->eng, vt_trace_closest(h->eng, static_cast<const vt_ray*>(d), n,
                                                 static_cast<vt_hit*>(out))))
    return nullptr;
  return ab;
}

// traceAny(handle, rays: Float32Array(8 n)) -> Uint32Array(n) of 0/1
napi_value TraceAny(napi_env env, napi_callback_info info) {
  auto a = args_of(env, info, 2);
  Handle* h = handle_of(env, a[0]);
  if (!h) return nullptr;
  bool is_ta = false;
  napi_is_typedarray(env, a[1], &is_ta);
  napi_typedarray_type t;
  size_t len = 0, off = 0;
  void* d = nullptr;
  napi_value ab0;
  if (!is_ta || napi_get_typedarray_info(env, a[1], &t, &len, &d, &ab0, &off) != napi_ok ||
      t != napi_float32_array || len % 8) {
    throw_type(env, "rays must be a Float32Array of 8 floats per ray");
    return nullptr;
  }
  const uint32_t n = (uint32_t)(len / 8);
  void* out = nullptr;
  napi_value ab, arr;
  NAPI_OK(napi_create_arraybuffer(env, (size_t)n * 4, &out, &ab));
  if (throw_status(env, h->eng, vt_trace_any(h->eng, static_cast<const vt_ray*>(d), n,
                                             static_cast<uint32_t*>(out))))
    return nullptr;
  NAPI_OK(napi_create_typedarray(env, napi_uint32_array, n, ab, 0, &arr));
  return arr;
}

// invertMat4(Float32Array(16)) -> Float32Array(16) | null
napi_value InvertMat4(napi_env env, napi_callback_info info) {
  auto a = args_of(env, info, 1);
  bool is_ta = false;
  napi_is_typedarray(env, a[0], &is_ta);
  napi_typedarray_type t;
  size_t len = 0, off = 0;
  void* d = nullptr;
  napi_value ab0;
  if (!is_ta || napi_get_typedarray_info(env, a[0], &t, &len, &d, &ab0, &off) != napi_ok ||
      t != napi_float32_array || len != 16) {
    throw_type(env, "expected a Float32Array(16)");
    return nullptr;
  }
  void* out = nullptr;
  napi_value ab, arr;
  NAPI_OK(napi_create_arraybuffer(env, 64, &out, &ab));
  if (vt_invert_mat4(static_cast<const float*>(d), static_cast<float*>(out)) != VT_OK) {
    napi_get_null(env, &arr);
    return arr;
  }
  NAPI_OK(napi_create_typedarray(env, napi_float32_array, 16, ab, 0, &arr));
  return arr;
}

// nrcConfig(handle, {enabled, trainPaths, learningRate, tableLearningRate, seed,
// teacherVertices}): the radiance cache of the path tracer (nrcEnabled /
// nrcConfig of the reference's WalkaroundGPUPipeline.ts:1860-1862).  A
// learningRate or seed re-initialises the cache's weights and Adam state;
// teacherVertices sets the training paths' suffix length; needs a scene.
napi_value NrcConfig(napi_env env, napi_callback_info info) {
  auto a = args_of(env, info, 2);
  Handle* h = handle_of(env, a[0]);
  if (!h) return nullptr;
  uint32_t enabled = 0, train_paths = 65536, seed = 1;
  float lr = 1e-2f, table_lr = 1e-1f;
  uint32_t teacher = 4;
  bool has_lr = false, has_seed = false, has_tlr = false, has_teacher = false;
  if (!is_nullish(env, a[1])) {
    if (!get_u32(env, a[1], "enabled", 0, &enabled)) return nullptr;
    if (!get_u32(env, a[1], "trainPaths", 65536, &train_paths)) return nullptr;
    napi_has_named_property(env, a[1], "learningRate", &has_lr);
    napi_has_named_property(env, a[1], "seed", &has_seed);
    if (!get_f32(env, a[1], "learningRate", 1e-2f, &lr)) return nullptr;
    if (!get_u32(env, a[1], "seed", 1, &seed)) return nullptr;
    napi_has_named_property(env, a[1], "tableLearningRate", &has_tlr);
    if (!get_f32(env, a[1], "tableLearningRate", 1e-1f, &table_lr)) return nullptr;
    napi_has_named_property(env, a[1], "teacherVertices", &has_teacher);
    if (!get_u32(env, a[1], "teacherVertices", 4, &teacher)) return nullptr;
  }
  if (has_lr || has_seed || has_tlr) {
    vt_nrc_options o{};
    o.learning_rate = lr;
    o.beta1 = 0.9f;
    o.beta2 = 0.999f;
    o.epsilon = 1e-8f;
    o.seed = seed;
    o.table_learning_rate = table_lr;
    if (throw_status(env, h->eng, vt_nrc_reset(h->eng, &o))) return nullptr;
  }
  if (throw_status(env, h->eng, vt_nrc_render_config(h->eng, enabled ? 1 : 0, train_paths))) return nullptr;
  if (has_teacher) throw_status(env, h->eng, vt_nrc_teacher_vertices(h->eng, teacher));
  return nullptr;
}

napi_value Version(napi_env env, napi_callback_info) {
  napi_value v;
  NAPI_OK(napi_create_string_utf8(env, vt_version(), NAPI_AUTO_LENGTH, &v));
  return v;
}

napi_value Init(napi_env env, napi_value exports) {
  const struct {
    const char* name;
    napi_callback cb;
  } fns[] = {{"create", Create},         {"destroy", Destroy},
             {"setScene", SetScene},     {"renderFrame", RenderFrame},
             {"capture", Capture},       {"reset", Reset},
             {"stats", Stats},           {"traceClosest", TraceClosest},
             {"traceAny", TraceAny},     {"invertMat4", InvertMat4},
             {"nrcConfig", NrcConfig},   {"version", Version},
             {"captureOutput", CaptureOutput}, {"captureAux", CaptureAux},
             {"renderFrames", RenderFrames}, {"refitPrimitive", RefitPrimitive}};
  for (const auto& f : fns) {
    napi_value fn;
    if (napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.cb, nullptr, &fn) != napi_ok) return nullptr;
    if (napi_set_named_property(env, exports, f.name, fn) != napi_ok) return nullptr;
  }
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
