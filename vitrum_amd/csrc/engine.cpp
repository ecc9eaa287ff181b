// engine.cpp -- the C ABI of libvthip.so (declared in include/vitrum_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/vitrum_hip.h"
#include "builder.h"
#include "device_layout.h"
#include "engine_internal.h"
#include "gpubuild.h"
#include "traverse.h"

namespace {

thread_local std::string g_global_error;

vt_status set_err(vt_engine* e, vt_status code, const std::string& msg) {
  if (e)
    e->last_error = msg;
  else
    g_global_error = msg;
  return code;
}

thread_local uint32_t g_fail_inject = 0;  // kInject* bits for calls without an engine

#define HIP_TRY(eng, expr)                                                               \
  do {                                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess)                                                                \
      return set_err((eng), VT_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

vt::PrimitiveIn to_prim(const vt_primitive& p) {
  return vt::PrimitiveIn{p.kind,      p.material,       p.positions,   p.position_floats,
                         p.normals,   p.normal_floats,  p.indices,     p.index_count,
                         p.transforms, p.transform_count};
}

vt_status check_prims(vt_engine* e, const vt_primitive* prims, uint32_t n) {
  if (n && !prims) return set_err(e, VT_ERR_INVALID_ARGUMENT, "primitives is NULL");
  for (uint32_t i = 0; i < n; ++i) {
    const vt_primitive& p = prims[i];
    if (p.kind > 1) return set_err(e, VT_ERR_INVALID_ARGUMENT, "unknown primitive kind");
    if (p.position_floats && !p.positions)
      return set_err(e, VT_ERR_INVALID_ARGUMENT, "primitive positions is NULL");
    if (p.normal_floats && !p.normals)
      return set_err(e, VT_ERR_INVALID_ARGUMENT, "primitive normals is NULL");
    if (p.index_count && !p.indices)
      return set_err(e, VT_ERR_INVALID_ARGUMENT, "primitive indices is NULL");
    if (p.transform_count && !p.transforms)
      return set_err(e, VT_ERR_INVALID_ARGUMENT, "primitive transforms is NULL");
    if (p.kind == VT_PRIM_MESH && p.transform_count > 1)
      return set_err(e, VT_ERR_INVALID_ARGUMENT, "a mesh takes at most one transform");
  }
  return VT_OK;
}

}  // namespace

vt_status engine_error(vt_engine* e, vt_status code, const std::string& msg) { return set_err(e, code, msg); }

uint32_t engine_build_threads(const vt_engine* e) {
  const uint32_t hw = std::max(1u, std::thread::hardware_concurrency());
  return e && e->build_threads ? std::min(e->build_threads, hw) : hw;
}

vt::BuildError engine_pack_scene(vt_engine* e, const vt::PrimitiveIn* prims, uint32_t count, vt::Pack& out) {
  if (!e->device_build_min) return vt::pack_scene(prims, count, engine_build_threads(e), out);
  const vt::BlasBuilder dev = [e](const std::vector<vt::BlasJobIn>& meshes, std::vector<vt::BlasOut>& outs) {
    return vtg::build_blas_forest(e, meshes, 4, 16, vt::kMaxBuildDepth, outs);
  };
  return vt::pack_scene(prims, count, engine_build_threads(e), out, &dev, e->device_build_min);
}

// ---------------------------------------------------------------------------
// Device buffer helpers
// ---------------------------------------------------------------------------

vt_status DeviceBuffer::ensure(vt_engine* e, size_t want) {
  if (want <= bytes) return VT_OK;
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
  if (want == 0) return VT_OK;
  hipError_t r = hipMalloc(&ptr, want);
  if (r != hipSuccess) {
    ptr = nullptr;
    return set_err(e, VT_ERR_OUT_OF_MEMORY,
                   "hipMalloc(" + std::to_string(want) + "): " + hipGetErrorString(r));
  }
  bytes = want;
  return VT_OK;
}

void DeviceBuffer::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

vt_status engine_upload(vt_engine* e, DeviceBuffer& dst, const void* src, size_t bytes) {
  vt_status s = dst.ensure(e, std::max<size_t>(bytes, 16));
  if (s) return s;
  if (bytes) HIP_TRY(e, hipMemcpyAsync(dst.ptr, src, bytes, hipMemcpyHostToDevice, e->stream));
  return VT_OK;
}

void SceneBuffers::release() {
  blas.release();
  tlas.release();
  blas4.release();
  tlas4.release();
  tris.release();
  instances.release();
  positions.release();
  normals.release();
  indices.release();
  tri_material.release();
  materials.release();
}

vt_status engine_upload_trace(vt_engine* e, const vt::DeviceLayout& layout, SceneBuffers& out) {
  vt_status s = VT_OK;
  s = s ? s : engine_upload(e, out.blas, layout.blas_nodes.data(), layout.blas_nodes.size() * 4);
  s = s ? s : engine_upload(e, out.tlas, layout.tlas_nodes.data(), layout.tlas_nodes.size() * 4);
  s = s ? s : engine_upload(e, out.blas4, layout.blas4_nodes.data(), layout.blas4_nodes.size() * 4);
  // A refit can collapse the TLAS differently; its Node128 tree never has
  // more nodes than the Node64 tree it comes from.
  s = s ? s : out.tlas4.ensure(e, std::max<size_t>((size_t)layout.tlas_node_count * vt::kNode128Words * 4, 16));
  s = s ? s : engine_upload(e, out.tlas4, layout.tlas4_nodes.data(), layout.tlas4_nodes.size() * 4);
  s = s ? s : engine_upload(e, out.tris, layout.tris.data(), layout.tris.size() * 4);
  s = s ? s : engine_upload(e, out.instances, layout.instances.data(), layout.instances.size() * 4);
  return s;
}

vtk::DeviceScene engine_device_scene(const SceneBuffers& b, const vt::DeviceLayout& L, bool* wide) {
  vtk::DeviceScene d{};
  d.blas_nodes = (const float4*)b.blas.ptr;
  d.tlas_nodes = (const float4*)b.tlas.ptr;
  d.tris = (const float4*)b.tris.ptr;
  d.instances = (const float4*)b.instances.ptr;
  d.tlas_root = L.tlas_root;
  d.tlas_node_count = L.tlas_node_count;
  d.tri_count = L.tri_count;
  d.inst_count = L.inst_count;
  d.blas4_nodes = (const float4*)b.blas4.ptr;
  d.tlas4_nodes = (const float4*)b.tlas4.ptr;
  d.tlas4_root = L.tlas4_root;
  if (wide) *wide = !L.blas4_nodes.empty() && L.stack4_need <= vtk::kStackCapacity;
  return d;
}

vtk::TraceParams engine_trace_params(vt_engine* e) {
  vtk::TraceParams p{};
  bool wide = false;
  p.scene = engine_device_scene(e->scene, *e->layout, &wide);
  p.wide = (wide && e->node_width == 4) ? 1u : 0u;
  p.counter = (uint32_t*)e->counter.ptr;
  p.error = (uint32_t*)e->error_word.ptr;
  p.spill = (uint2*)e->spill.ptr;
  p.refill_min = e->refill_min;
  p.steps = e->trace_steps;
  p.near_first = e->near_first;
  p.defer_min = e->defer_min;
  p.sided = e->hit_filter;
  return p;
}

// ---------------------------------------------------------------------------
// Lifecycle
// ---------------------------------------------------------------------------

// Packs, builds and uploads the traversal buffers of one scene (no shading
// data); used by the batched launch.
vt_status engine_build_trace_scene(vt_engine* eng, const vt_scene_desc* scene, SceneBuffers& out,
                                   vt::DeviceLayout& layout) {
  if (!scene) return set_err(eng, VT_ERR_INVALID_ARGUMENT, "scene is NULL");
  if (vt_status s = check_prims(eng, scene->primitives, scene->primitive_count)) return s;
  std::vector<vt::PrimitiveIn> prims(scene->primitive_count);
  for (uint32_t i = 0; i < scene->primitive_count; ++i) prims[i] = to_prim(scene->primitives[i]);
  vt::Pack pack;
  vt::BuildError be = engine_pack_scene(eng, prims.data(), (uint32_t)prims.size(), pack);
  if (be) return set_err(eng, be.code, be.message);
  std::string le = vt::build_device_layout(pack, layout);
  if (!le.empty()) return set_err(eng, VT_ERR_RANGE, le);
  HIP_TRY(eng, hipSetDevice(eng->device));
  vt_status s = engine_upload_trace(eng, layout, out);
  if (!s) {
    hipError_t r = hipStreamSynchronize(eng->stream);
    if (r != hipSuccess) s = set_err(eng, VT_ERR_HIP, hipGetErrorString(r));
  }
  if (s) out.release();
  return s;
}

extern "C" {

vt_status vt_invert_mat4(const float m[16], float out[16]) {
  return vt_guarded(nullptr, [&]() -> vt_status {
    if (!m || !out) return set_err(nullptr, VT_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!vt::invert_mat4(m, out)) return set_err(nullptr, VT_ERR_RANGE, "matrix is not invertible");
    return VT_OK;
  });
}

vt_status vt_bake_procedural_sky(const vt_environment* env, uint32_t width, uint32_t height, float* rgb_out,
                                 double* luminance_integral) {
  return vt_guarded(nullptr, [&]() -> vt_status {
    if (!env || !rgb_out || !luminance_integral) return set_err(nullptr, VT_ERR_INVALID_ARGUMENT, "NULL argument");
    if (env->kind != VT_ENV_PROCEDURAL_SKY)
      return set_err(nullptr, VT_ERR_INVALID_ARGUMENT, "the environment is not a procedural sky");
    std::vector<float> rgb;
    double integral = 0;
    const std::string err = engine_bake_sky(*env, width, height, rgb, &integral);
    if (!err.empty()) return set_err(nullptr, VT_ERR_RANGE, err);
    std::memcpy(rgb_out, rgb.data(), rgb.size() * sizeof(float));
    *luminance_integral = integral;
    return VT_OK;
  });
}

const char* vt_version(void) { return "vitrum-hip 0.1.0 gfx950"; }

vt_status vt_last_error(const vt_engine* eng, char* buf, size_t len) {
  if (!buf || len == 0) return VT_ERR_INVALID_ARGUMENT;
  const std::string& m = eng ? eng->last_error : g_global_error;
  std::snprintf(buf, len, "%s", m.c_str());
  return VT_OK;
}

vt_status vt_create(int32_t device_ordinal, const vt_options* opts, vt_engine** out) {
  return vt_guarded(nullptr, [&]() -> vt_status {
    if (!out) return set_err(nullptr, VT_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
      return set_err(nullptr, VT_ERR_NO_DEVICE, "no HIP device is available");
    if (device_ordinal < 0 || device_ordinal >= count)
      return set_err(nullptr, VT_ERR_RANGE, "device ordinal out of range");
    if (opts && opts->sampling > 1)
      return set_err(nullptr, VT_ERR_RANGE, "sampling must be 0 (pcg) or 1 (sobol)");
    if (opts && opts->light_selection > 1)
      return set_err(nullptr, VT_ERR_RANGE, "light_selection must be 0 (light tree) or 1 (uniform)");
    if (opts && opts->bvh_traversal > 1)
      return set_err(nullptr, VT_ERR_RANGE, "bvh_traversal must be 0 (binary) or 1 (cwbvh-closest)");
    std::unique_ptr<vt_engine> e(new vt_engine());
    e->device = device_ordinal;
    if (opts) e->opts = *opts;
    if (e->opts.max_bounces == 0) e->opts.max_bounces = 8;
    HIP_TRY(nullptr, hipSetDevice(device_ordinal));
    HIP_TRY(nullptr, hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    HIP_TRY(nullptr, hipEventCreate(&e->ev_start));
    HIP_TRY(nullptr, hipEventCreate(&e->ev_stop));
    hipDeviceProp_t prop;
    HIP_TRY(nullptr, hipGetDeviceProperties(&prop, device_ordinal));
    e->cu_count = (uint32_t)prop.multiProcessorCount;
    int per_cu = 0;
    HIP_TRY(nullptr, vtk::trace_occupancy(false, &per_cu));
    e->trace_blocks_per_cu = std::max(1, per_cu);
    e->trace_grid = e->cu_count * e->trace_blocks_per_cu;
    vt_status s = e->counter.ensure(e.get(), 64);
    if (!s) s = e->error_word.ensure(e.get(), 64);
    const size_t spill = (size_t)(vtk::kStackCapacity - vtk::kLdsDepth) * e->trace_grid *
                         vtk::kTraceBlock * sizeof(uint2);
    if (!s) s = e->spill.ensure(e.get(), spill);
    if (s) return set_err(nullptr, s, e->last_error);
    HIP_TRY(nullptr, vtk::launch_zero((uint32_t*)e->error_word.ptr, 16, e->stream));
    HIP_TRY(nullptr, hipStreamSynchronize(e->stream));
    if (engine_noalloc_setup(e.get())) return set_err(nullptr, VT_ERR_HIP, e->last_error);
    *out = e.release();
    return VT_OK;
  });
}

vt_status vt_destroy(vt_engine* eng) {
  return vt_guarded(eng, [&]() -> vt_status {
    if (!eng) return VT_OK;
    (void)hipSetDevice(eng->device);
    (void)hipStreamSynchronize(eng->stream);
    eng->scene.release();
    eng->counter.release();
    eng->error_word.release();
    eng->spill.release();
    eng->tmp_rays.release();
    eng->tmp_out.release();
    eng->tmp_counts.release();
    engine_release_cwbvh(eng);
    engine_release_refit(eng);
    engine_release_gpubuild(eng);
    for (auto& kv : eng->user_allocs) (void)hipFree(kv);
    engine_release_render(eng);
    engine_release_nrc(eng);
    (void)hipEventDestroy(eng->ev_start);
    (void)hipEventDestroy(eng->ev_stop);
    (void)hipStreamDestroy(eng->stream);
    delete eng;
    return VT_OK;
  });
}

vt_status vt_set_scene(vt_engine* eng, const vt_scene_desc* scene) {
  return vt_guarded(eng, [&]() -> vt_status {
    if (!eng) return set_err(nullptr, VT_ERR_INVALID_ARGUMENT, "engine is NULL");
    if (!scene) return set_err(eng, VT_ERR_INVALID_ARGUMENT, "scene is NULL");
    if (vt_status s = check_prims(eng, scene->primitives, scene->primitive_count)) return s;
    if (scene->material_count && !scene->materials)
      return set_err(eng, VT_ERR_INVALID_ARGUMENT, "materials is NULL");
    for (uint32_t i = 0; i < scene->primitive_count; ++i)
      if (scene->primitives[i].material >= std::max<uint32_t>(1, scene->material_count) &&
          scene->material_count > 0)
        return set_err(eng, VT_ERR_RANGE, "primitive material index out of range");
    for (uint32_t i = 0; i < scene->material_count; ++i) {
      // Le = emissive x emissiveIntensity, rounded once to f32, must be finite
      // and non-negative (an f32 overflow from finite JavaScript numbers would
      // otherwise reach the light tree as an infinite power).
      const vt_material& m = scene->materials[i];
      const float ext[] = {m.clearcoat, m.clearcoat_roughness, m.sheen, m.sheen_roughness, m.sheen_color[0],
                           m.sheen_color[1], m.sheen_color[2], m.anisotropy, m.anisotropy_rotation,
                           m.specular_color[0], m.specular_color[1], m.specular_color[2], m.specular_intensity};
      for (float x : ext)
        if (!std::isfinite(x))
          return set_err(eng, VT_ERR_RANGE, "material " + std::to_string(i) + ": extension lobe parameters must be finite");
      for (int k = 0; k < 3; ++k) {
        const float le = (float)((double)m.emissive[k] * (double)m.emissive_intensity);
        if (!std::isfinite(le) || le < 0.0f)
          return set_err(eng, VT_ERR_RANGE, "material " + std::to_string(i) +
                                                ": emissive x emissiveIntensity must be finite and >= 0 in f32");
      }
    }
    if (eng->fail_inject & kInjectPackThrow) {
      eng->fail_inject &= ~kInjectPackThrow;
      throw std::bad_alloc();
    }
    std::vector<vt::PrimitiveIn> prims(scene->primitive_count);
    for (uint32_t i = 0; i < scene->primitive_count; ++i) prims[i] = to_prim(scene->primitives[i]);
    auto pack = std::make_unique<vt::Pack>();
    vt::BuildError be = engine_pack_scene(eng, prims.data(), (uint32_t)prims.size(), *pack);
    if (be) return set_err(eng, be.code, be.message);
    auto layout = std::make_unique<vt::DeviceLayout>();
    std::string le = vt::build_device_layout(*pack, *layout);
    if (!le.empty()) return set_err(eng, VT_ERR_RANGE, le);
    // Single-sided triangles (materialAcceptsSidedHit, material.wgsl.ts:
    // 3259-3263): opaque materials without doubleSided (the default) have no
    // back-face surfaces for the path tracer's closest-hit and visibility rays.
    for (uint32_t t = 0; t < layout->tri_count; ++t) {
      const uint32_t mid = pack->tri_material[t];
      const bool single = mid < scene->material_count ? scene->materials[mid].double_sided == 0 : true;
      layout->tris[(size_t)t * vt::kTriWords + 9] = single ? 1.0f : 0.0f;
    }
    std::vector<vt_material> materials(scene->materials, scene->materials + scene->material_count);
    HIP_TRY(eng, hipSetDevice(eng->device));
    // Build and upload everything into fresh objects -- traversal buffers,
    // then the shading data -- and bind them only when all of it succeeded:
    // any failure leaves the previous scene bound and usable.
    SceneBuffers nb;
    RenderSceneData* rd = nullptr;
    struct Unbound {  // frees what was not bound, also when an exception unwinds
      SceneBuffers& nb;
      RenderSceneData*& rd;
      ~Unbound() {
        nb.release();
        engine_discard_render(rd);
      }
    } unbound{nb, rd};
    vt_status s = engine_upload_trace(eng, *layout, nb);
    s = s ? s : engine_upload(eng, nb.positions, pack->positions.data(), pack->positions.size() * 4);
    s = s ? s : engine_upload(eng, nb.normals, pack->normals.data(), pack->normals.size() * 4);
    s = s ? s : engine_upload(eng, nb.indices, pack->indices.data(), pack->indices.size() * 4);
    s = s ? s : engine_upload(eng, nb.tri_material, pack->tri_material.data(), pack->tri_material.size() * 4);
    s = s ? s : engine_prepare_render(eng, *pack, materials, scene->environment, scene->scene_radius,
                                      scene->emitters, scene->emitter_count, &rd);
    if (!s) {
      hipError_t r = hipStreamSynchronize(eng->stream);
      if (r != hipSuccess) s = set_err(eng, VT_ERR_HIP, hipGetErrorString(r));
    }
    if (s) return s;
    // Commit: nothing below can fail.
    engine_release_cwbvh(eng);
    engine_refit_scene_changed(eng);
    ++eng->scene_generation;
    eng->scene.release();
    eng->scene = nb;
    nb = SceneBuffers();
    engine_commit_render(eng, rd);
    rd = nullptr;
    eng->layout = std::move(layout);
    eng->pack = std::move(pack);
    eng->materials = std::move(materials);
    eng->scene_radius = scene->scene_radius;
    eng->has_scene = true;
    eng->samples_accumulated = 0;
    return VT_OK;
  });
}

vt_status vt_set_build_threads(vt_engine* eng, uint32_t threads) {
  if (!eng) return set_err(nullptr, VT_ERR_INVALID_ARGUMENT, "engine is NULL");
  eng->build_threads = threads;
  return VT_OK;
}

vt_status vt_set_device_build(vt_engine* eng, uint32_t min_triangles) {
  if (!eng) return set_err(nullptr, VT_ERR_INVALID_ARGUMENT, "engine is NULL");
  eng->device_build_min = min_triangles;
  return VT_OK;
}

vt_status vt_debug_inject_failure(vt_engine* eng, uint32_t flags) {
  if (eng)
    eng->fail_inject = flags;
  else
    g_fail_inject = flags;
  return VT_OK;
}

vt_status vt_device_alloc(vt_engine* eng, size_t bytes, void** d_ptr) {
  return vt_guarded(eng, [&]() -> vt_status {
    if (!eng || !d_ptr) return set_err(eng, VT_ERR_INVALID_ARGUMENT, "NULL argument");
    HIP_TRY(eng, hipSetDevice(eng->device));
    HIP_TRY(eng, hipMalloc(d_ptr, std::max<size_t>(bytes, 16)));
    eng->user_allocs.push_back(*d_ptr);
    return VT_OK;
  });
}

vt_status vt_device_free(vt_engine* eng, void* d_ptr) {
  return vt_guarded(eng, [&]() -> vt_status {
    if (!eng) return set_err(nullptr, VT_ERR_INVALID_ARGUMENT, "engine is NULL");
    auto it = std::find(eng->user_allocs.begin(), eng->user_allocs.end(), d_ptr);
    if (it == eng->user_allocs.end())
      return set_err(eng, VT_ERR_INVALID_ARGUMENT, "pointer was not allocated by vt_device_alloc");
    eng->user_allocs.erase(it);
    HIP_TRY(eng, hipStreamSynchronize(eng->stream));
    HIP_TRY(eng, hipFree(d_ptr));
    return VT_OK;
  });
}

vt_status vt_memcpy_h2d(vt_engine* eng, void* d_dst, const void* src, size_t bytes) {
  return vt_guarded(eng, [&]() -> vt_status {
    if (!eng || (!src && bytes) || (!d_dst && bytes))
      return set_err(eng, VT_ERR_INVALID_ARGUMENT, "NULL argument");
    HIP_TRY(eng, hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, eng->stream));
    HIP_TRY(eng, hipStreamSynchronize(eng->stream));
    return VT_OK;
  });
}

vt_status vt_memcpy_d2h(vt_engine* eng, void* dst, const void* d_src, size_t bytes) {
  return vt_guarded(eng, [&]() -> vt_status {
    if (!eng || (!dst && bytes) || (!d_src && bytes))
      return set_err(eng, VT_ERR_INVALID_ARGUMENT, "NULL argument");
    HIP_TRY(eng, hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, eng->stream));
    HIP_TRY(eng, hipStreamSynchronize(eng->stream));
    return VT_OK;
  });
}

vt_status vt_memcpy_d2d(vt_engine* eng, void* d_dst, const void* d_src, size_t bytes) {
  return vt_guarded(eng, [&]() -> vt_status {
    if (!eng || (!d_dst && bytes) || (!d_src && bytes))
      return set_err(eng, VT_ERR_INVALID_ARGUMENT, "NULL argument");
    if ((bytes | (uintptr_t)d_dst | (uintptr_t)d_src) & 15u)
      return set_err(eng, VT_ERR_INVALID_ARGUMENT, "pointers and size must be multiples of 16 bytes");
    HIP_TRY(eng, hipSetDevice(eng->device));
    HIP_TRY(eng, vtk::launch_copy16((uint4*)d_dst, (const uint4*)d_src, bytes / 16, eng->stream));
    HIP_TRY(eng, hipStreamSynchronize(eng->stream));
    return VT_OK;
  });
}

vt_status vt_synchronize(vt_engine* eng) {
  return vt_guarded(eng, [&]() -> vt_status {
    if (!eng) return set_err(nullptr, VT_ERR_INVALID_ARGUMENT, "engine is NULL");
    if (vt_status s = engine_nrc_join(eng)) return s;
    HIP_TRY(eng, hipStreamSynchronize(eng->stream));
    return VT_OK;
  });
}

vt_status vt_get_stream(vt_engine* eng, void** stream) {
  return vt_guarded(eng, [&]() -> vt_status {
    if (!eng || !stream) return set_err(eng, VT_ERR_INVALID_ARGUMENT, "NULL argument");
    *stream = (void*)eng->stream;
    return VT_OK;
  });
}

vt_status vt_stats(vt_engine* eng, vt_frame_stats* out) {
  return vt_guarded(eng, [&]() -> vt_status {
    if (!eng || !out) return set_err(eng, VT_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = vt_frame_stats{};
    if (eng->trace_timed) {
      HIP_TRY(eng, hipEventSynchronize(eng->ev_stop));
      float ms = 0;
      HIP_TRY(eng, hipEventElapsedTime(&ms, eng->ev_start, eng->ev_stop));
      out->last_trace_ms = ms;
    }
    out->last_frame_ms = eng->last_frame_ms;
    out->samples_accumulated = eng->samples_accumulated;
    out->last_ray_count = eng->last_ray_count;
    out->device_bytes = eng->scene.blas.bytes + eng->scene.tlas.bytes + eng->scene.tris.bytes +
                        eng->scene.instances.bytes + eng->scene.positions.bytes +
                        eng->scene.normals.bytes + eng->scene.indices.bytes;
    return VT_OK;
  });
}

// ---------------------------------------------------------------------------
// Traversal
// ---------------------------------------------------------------------------

static vt_status trace_device(vt_engine* eng, const vt_ray* d_rays, uint32_t n, void* d_out,
                              bool any, uint4* d_counts) {
  if (!eng) return set_err(nullptr, VT_ERR_INVALID_ARGUMENT, "engine is NULL");
  if (!eng->has_scene) return set_err(eng, VT_ERR_STATE, "no scene is set");
  if (n && (!d_rays || !d_out)) return set_err(eng, VT_ERR_INVALID_ARGUMENT, "NULL buffer");
  if (n > 0x7FFFFFFFu) return set_err(eng, VT_ERR_RANGE, "too many rays");
  HIP_TRY(eng, hipSetDevice(eng->device));
  vtk::TraceParams p = engine_trace_params(eng);
  p.rays = (const float4*)d_rays;
  p.hits = any ? nullptr : (uint32_t*)d_out;
  p.occluded = any ? (uint32_t*)d_out : nullptr;
  p.counts = d_counts;
  p.n = n;
  eng->last_ray_count = n;
  if (n == 0) return VT_OK;
  // The cwbvh-closest mode covers closest hits; any-hit rays and the work
  // counters (defined on the binary order) stay on the binary kernel.
  if (!any && !d_counts && eng->opts.bvh_traversal == 1) {
    HIP_TRY(eng, hipEventRecord(eng->ev_start, eng->stream));
    if (vt_status s = engine_trace_cwbvh(eng, (const float4*)d_rays, n, nullptr, (uint32_t*)d_out, nullptr,
                                         eng->hit_filter))
      return s;
    HIP_TRY(eng, hipEventRecord(eng->ev_stop, eng->stream));
    eng->trace_timed = true;
    return VT_OK;
  }
  HIP_TRY(eng, hipEventRecord(eng->ev_start, eng->stream));
  HIP_TRY(eng, vtk::launch_trace(p, any, d_counts != nullptr, eng->trace_grid, eng->stream));
  HIP_TRY(eng, hipEventRecord(eng->ev_stop, eng->stream));
  eng->trace_timed = true;
  return VT_OK;
}

static vt_status check_error_word(vt_engine* eng) {
  uint32_t err = 0;
  HIP_TRY(eng, hipMemcpyAsync(&err, eng->error_word.ptr, 4, hipMemcpyDeviceToHost, eng->stream));
  HIP_TRY(eng, hipStreamSynchronize(eng->stream));
  if (err) {
    (void)vtk::launch_zero((uint32_t*)eng->error_word.ptr, 1, eng->stream);
    return set_err(eng, VT_ERR_HIP, "traversal stack capacity exceeded");
  }
  return VT_OK;
}

static vt_status trace_host(vt_engine* eng, const vt_ray* rays, uint32_t n, void* out,
                            size_t out_stride, bool any, uint32_t* counts) {
  if (!eng) return set_err(nullptr, VT_ERR_INVALID_ARGUMENT, "engine is NULL");
  if (n && (!rays || !out)) return set_err(eng, VT_ERR_INVALID_ARGUMENT, "NULL buffer");
  if (n == 0) return VT_OK;
  HIP_TRY(eng, hipSetDevice(eng->device));
  vt_status s = eng->tmp_rays.ensure(eng, (size_t)n * sizeof(vt_ray));
  if (!s) s = eng->tmp_out.ensure(eng, (size_t)n * out_stride);
  if (!s && counts) s = eng->tmp_counts.ensure(eng, (size_t)n * 16);
  if (s) return s;
  HIP_TRY(eng, hipMemcpyAsync(eng->tmp_rays.ptr, rays, (size_t)n * sizeof(vt_ray),
                              hipMemcpyHostToDevice, eng->stream));
  s = trace_device(eng, (const vt_ray*)eng->tmp_rays.ptr, n, eng->tmp_out.ptr, any,
                   counts ? (uint4*)eng->tmp_counts.ptr : nullptr);
  if (s) return s;
  HIP_TRY(eng, hipMemcpyAsync(out, eng->tmp_out.ptr, (size_t)n * out_stride, hipMemcpyDeviceToHost,
                              eng->stream));
  if (counts)
    HIP_TRY(eng, hipMemcpyAsync(counts, eng->tmp_counts.ptr, (size_t)n * 16, hipMemcpyDeviceToHost,
                                eng->stream));
  HIP_TRY(eng, hipStreamSynchronize(eng->stream));
  return check_error_word(eng);
}

vt_status vt_trace_closest(vt_engine* eng, const vt_ray* rays, uint32_t n, vt_hit* hits) {
  return vt_guarded(eng, [&]() -> vt_status {
    return trace_host(eng, rays, n, hits, sizeof(vt_hit), false, nullptr);
  });
}

vt_status vt_trace_any(vt_engine* eng, const vt_ray* rays, uint32_t n, uint32_t* occluded) {
  return vt_guarded(eng, [&]() -> vt_status {
    return trace_host(eng, rays, n, occluded, 4, true, nullptr);
  });
}

vt_status vt_trace_closest_counts(vt_engine* eng, const vt_ray* rays, uint32_t n, vt_hit* hits,
                                  uint32_t* counts) {
  return vt_guarded(eng, [&]() -> vt_status {
    if (!counts) return set_err(eng, VT_ERR_INVALID_ARGUMENT, "counts is NULL");
    return trace_host(eng, rays, n, hits, sizeof(vt_hit), false, counts);
  });
}

vt_status vt_set_counting(vt_engine* eng, int32_t enabled) {
  return vt_guarded(eng, [&]() -> vt_status {
    if (!eng) return set_err(nullptr, VT_ERR_INVALID_ARGUMENT, "engine is NULL");
    eng->counting = enabled != 0;
    return VT_OK;
  });
}

vt_status vt_trace_closest_device(vt_engine* eng, const vt_ray* d_rays, uint32_t n, vt_hit* d_hits) {
  return vt_guarded(eng, [&]() -> vt_status {
    return trace_device(eng, d_rays, n, d_hits, false, nullptr);
  });
}

vt_status vt_trace_any_device(vt_engine* eng, const vt_ray* d_rays, uint32_t n, uint32_t* d_occluded) {
  return vt_guarded(eng, [&]() -> vt_status {
    return trace_device(eng, d_rays, n, d_occluded, true, nullptr);
  });
}

vt_status vt_set_shade_sort(vt_engine* eng, uint32_t enabled) {
  return vt_guarded(eng, [&]() -> vt_status {
    if (!eng) return set_err(nullptr, VT_ERR_INVALID_ARGUMENT, "engine is NULL");
    if (enabled > 1) return set_err(eng, VT_ERR_RANGE, "enabled must be 0 or 1");
    eng->shade_sort = enabled;
    return VT_OK;
  });
}

vt_status vt_set_trace_tuning(vt_engine* eng, uint32_t refill_min, uint32_t steps, uint32_t near_first,
                              uint32_t defer_min) {
  return vt_guarded(eng, [&]() -> vt_status {
    if (!eng) return set_err(nullptr, VT_ERR_INVALID_ARGUMENT, "engine is NULL");
    if (refill_min < 1 || refill_min > 64 || steps < 1 || near_first > 2 || defer_min < 1 || defer_min > 64)
      return set_err(eng, VT_ERR_RANGE,
                     "refill_min and defer_min must be in [1, 64], steps >= 1 and near_first 0, 1 or 2");
    eng->refill_min = refill_min;
    eng->trace_steps = steps;
    eng->near_first = near_first;
    eng->defer_min = defer_min;
    return VT_OK;
  });
}

vt_status vt_selftest_fp32(vt_engine* eng, const float* a, const float* b, float* q, float* r,
                           uint32_t n) {
  return vt_guarded(eng, [&]() -> vt_status {
    if (!eng || (n && (!a || !b || !q || !r)))
      return set_err(eng, VT_ERR_INVALID_ARGUMENT, "NULL argument");
    if (n == 0) return VT_OK;
    HIP_TRY(eng, hipSetDevice(eng->device));
    DeviceBuffer buf;
    if (vt_status s = buf.ensure(eng, (size_t)n * 16)) return s;
    float* d = (float*)buf.ptr;
    HIP_TRY(eng, hipMemcpyAsync(d, a, (size_t)n * 4, hipMemcpyHostToDevice, eng->stream));
    HIP_TRY(eng, hipMemcpyAsync(d + n, b, (size_t)n * 4, hipMemcpyHostToDevice, eng->stream));
    HIP_TRY(eng, vtk::launch_fp32_selftest(d, d + n, d + 2 * (size_t)n, d + 3 * (size_t)n, n, eng->stream));
    HIP_TRY(eng, hipMemcpyAsync(q, d + 2 * (size_t)n, (size_t)n * 4, hipMemcpyDeviceToHost, eng->stream));
    HIP_TRY(eng, hipMemcpyAsync(r, d + 3 * (size_t)n, (size_t)n * 4, hipMemcpyDeviceToHost, eng->stream));
    HIP_TRY(eng, hipStreamSynchronize(eng->stream));
    buf.release();
    return VT_OK;
  });
}

// ---------------------------------------------------------------------------
// Host-side builders
// ---------------------------------------------------------------------------

vt_status vt_build_blas(const float* positions, uint32_t position_floats, uint32_t position_stride,
                        const uint32_t* indices, uint32_t index_words, uint32_t index_stride,
                        const uint32_t* tri_material_ids, uint32_t max_leaf, uint32_t num_bins,
                        uint32_t max_depth, uint32_t* out_node_words, uint32_t* out_node_count,
                        uint32_t* out_reordered_indices, uint32_t* out_reordered_mat,
                        uint32_t* out_reordered_to_source) {
  return vt_guarded(nullptr, [&]() -> vt_status {
    if ((position_floats && !positions) || (index_words && !indices) || !out_node_words ||
        !out_node_count || (index_words && !out_reordered_indices))
      return set_err(nullptr, VT_ERR_INVALID_ARGUMENT, "NULL argument");
    vt::BlasOut o;
    vt::BuildError e = vt::build_blas(positions, position_floats, position_stride, indices, index_words,
                                      index_stride, tri_material_ids, max_leaf, num_bins, max_depth, o);
    if (e) return set_err(nullptr, e.code, e.message);
    std::memcpy(out_node_words, o.nodes.data(), o.nodes.size() * 4);
    *out_node_count = (uint32_t)(o.nodes.size() / 8);
    if (!o.reordered_indices.empty())
      std::memcpy(out_reordered_indices, o.reordered_indices.data(), o.reordered_indices.size() * 4);
    if (out_reordered_mat && !o.reordered_mat.empty())
      std::memcpy(out_reordered_mat, o.reordered_mat.data(), o.reordered_mat.size() * 4);
    if (out_reordered_to_source && !o.reordered_to_source.empty())
      std::memcpy(out_reordered_to_source, o.reordered_to_source.data(),
                  o.reordered_to_source.size() * 4);
    return VT_OK;
  });
}

vt_status vt_build_tlas(uint32_t n, const uint32_t* blas_ids, const double* aabb_min,
                        const double* aabb_max, const float* world_to_local, uint32_t max_leaf,
                        uint32_t num_bins, uint32_t max_depth, uint32_t* out_nodes,
                        uint32_t* out_node_count, uint32_t* out_instance_indices,
                        uint32_t* out_blas_roots, float* out_transforms) {
  return vt_guarded(nullptr, [&]() -> vt_status {
    if (n && (!blas_ids || !aabb_min || !aabb_max || !world_to_local))
      return set_err(nullptr, VT_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!out_nodes || !out_node_count || !out_instance_indices || !out_blas_roots || !out_transforms)
      return set_err(nullptr, VT_ERR_INVALID_ARGUMENT, "NULL output");
    vt::TlasOut o;
    vt::BuildError e = vt::build_tlas(n, blas_ids, aabb_min, aabb_max, world_to_local, max_leaf,
                                      num_bins, max_depth, o);
    if (e) return set_err(nullptr, e.code, e.message);
    std::memcpy(out_nodes, o.nodes.data(), o.nodes.size() * 4);
    *out_node_count = (uint32_t)(o.nodes.size() / 8);
    std::memcpy(out_instance_indices, o.instance_indices.data(), o.instance_indices.size() * 4);
    std::memcpy(out_blas_roots, o.blas_roots.data(), o.blas_roots.size() * 4);
    std::memcpy(out_transforms, o.transforms.data(), o.transforms.size() * 4);
    return VT_OK;
  });
}

struct vt_pack {
  vt::Pack p;
};

vt_status vt_pack_scene(const vt_primitive* prims, uint32_t prim_count, uint32_t threads,
                        vt_pack** out) {
  return vt_guarded(nullptr, [&]() -> vt_status {
    if (!out) return set_err(nullptr, VT_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (vt_status s = check_prims(nullptr, prims, prim_count)) return s;
    if (g_fail_inject & kInjectPackThrow) {
      g_fail_inject &= ~kInjectPackThrow;
      throw std::bad_alloc();
    }
    std::vector<vt::PrimitiveIn> v(prim_count);
    for (uint32_t i = 0; i < prim_count; ++i) v[i] = to_prim(prims[i]);
    auto pk = std::make_unique<vt_pack>();
    vt::BuildError e = vt::pack_scene(v.data(), prim_count, threads, pk->p);
    if (e) return set_err(nullptr, e.code, e.message);
    *out = pk.release();
    return VT_OK;
  });
}

static const void* pack_field(const vt::Pack& p, int32_t f, size_t* words) {
  auto ret = [&](const auto& vec) -> const void* {
    *words = vec.size();
    return vec.data();
  };
  switch (f) {
    case VT_PACK_POSITIONS: return ret(p.positions);
    case VT_PACK_NORMALS: return ret(p.normals);
    case VT_PACK_INDICES: return ret(p.indices);
    case VT_PACK_TRI_MATERIAL: return ret(p.tri_material);
    case VT_PACK_TRI_SOURCE: return ret(p.tri_source);
    case VT_PACK_TRI_PRIMITIVE: return ret(p.tri_primitive);
    case VT_PACK_BVH_NODES: return ret(p.bvh_nodes);
    case VT_PACK_TLAS_NODES: return ret(p.tlas_nodes);
    case VT_PACK_TLAS_INSTANCE_INDICES: return ret(p.tlas_instance_indices);
    case VT_PACK_TLAS_BLAS_ROOTS: return ret(p.tlas_blas_roots);
    case VT_PACK_TLAS_W2L: return ret(p.tlas_w2l);
    case VT_PACK_TLAS_L2W: return ret(p.tlas_l2w);
    case VT_PACK_INSTANCE_PRIMITIVE: return ret(p.instance_primitive);
    case VT_PACK_INSTANCE_SOURCE: return ret(p.instance_source);
    default: *words = 0; return nullptr;
  }
}

int64_t vt_pack_field_words(const vt_pack* p, int32_t field) {
  if (!p || field < 0 || field >= VT_PACK_FIELD_COUNT) return -1;
  size_t w = 0;
  pack_field(p->p, field, &w);
  return (int64_t)w;
}

vt_status vt_pack_copy_field(const vt_pack* p, int32_t field, void* dst, uint64_t dst_words) {
  return vt_guarded(nullptr, [&]() -> vt_status {
    if (!p || field < 0 || field >= VT_PACK_FIELD_COUNT)
      return set_err(nullptr, VT_ERR_INVALID_ARGUMENT, "bad pack or field");
    size_t w = 0;
    const void* src = pack_field(p->p, field, &w);
    if (w > dst_words) return set_err(nullptr, VT_ERR_RANGE, "destination too small");
    if (w && !dst) return set_err(nullptr, VT_ERR_INVALID_ARGUMENT, "dst is NULL");
    if (w) std::memcpy(dst, src, w * 4);
    return VT_OK;
  });
}

void vt_pack_free(vt_pack* p) { delete p; }

vt_status vt_scene_copy_field(vt_engine* eng, int32_t field, void* dst, uint64_t dst_words, uint64_t* words) {
  return vt_guarded(eng, [&]() -> vt_status {
    if (!eng) return VT_ERR_INVALID_ARGUMENT;
    if (!eng->has_scene || !eng->pack || !eng->layout) return set_err(eng, VT_ERR_STATE, "no scene is set");
    const DeviceBuffer* dev = nullptr;
    size_t w = 0;
    const void* src = nullptr;
    switch (field) {
      case VT_SCENE_BLAS_NODE64:
        dev = &eng->scene.blas, w = eng->layout->blas_nodes.size();
        break;
      case VT_SCENE_BLAS_NODE128:
        dev = &eng->scene.blas4, w = eng->layout->blas4_nodes.size();
        break;
      case VT_SCENE_TRIANGLES:
        dev = &eng->scene.tris, w = eng->layout->tris.size();
        break;
      default:
        if (field < 0 || field >= VT_PACK_FIELD_COUNT) return set_err(eng, VT_ERR_INVALID_ARGUMENT, "bad field");
        src = pack_field(*eng->pack, field, &w);
    }
    if (words) *words = w;
    if (!dst) return VT_OK;
    if (w > dst_words) return set_err(eng, VT_ERR_RANGE, "destination too small");
    if (!w) return VT_OK;
    if (dev) {
      HIP_TRY(eng, hipSetDevice(eng->device));
      HIP_TRY(eng, hipMemcpyAsync(dst, dev->ptr, w * 4, hipMemcpyDeviceToHost, eng->stream));
      HIP_TRY(eng, hipStreamSynchronize(eng->stream));
    } else {
      std::memcpy(dst, src, w * 4);
    }
    return VT_OK;
  });
}

}  // extern "C"

vt_status vt_set_node_width(vt_engine* eng, uint32_t width) {
  return vt_guarded(eng, [&]() -> vt_status {
    if (!eng) return set_err(nullptr, VT_ERR_INVALID_ARGUMENT, "engine is NULL");
    if (width != 2 && width != 4) return set_err(eng, VT_ERR_RANGE, "node width must be 2 or 4");
    eng->node_width = width;
    return VT_OK;
  });
}

static_assert(vt::kWideKids == vtk::kWideKids && vt::kNode128Words == 4 * vtk::kWideNodeFloat4s,
              "the wide-node layout of the builder and of the traversal kernel differ");

vt_status vt_set_hit_filter(vt_engine* eng, uint32_t filter) {
  return vt_guarded(eng, [&]() -> vt_status {
    if (!eng) return set_err(nullptr, VT_ERR_INVALID_ARGUMENT, "engine is NULL");
    if (filter > 1) return set_err(eng, VT_ERR_RANGE, "hit filter must be 0 (raw) or 1 (sided wrappers)");
    eng->hit_filter = filter;
    return VT_OK;
  });
}
