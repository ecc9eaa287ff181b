// engine_internal.h -- the vt_engine object behind the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <exception>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/vitrum_hip.h"
#include "builder.h"
#include "device_layout.h"
#include "traverse.h"

// Records `msg` as the engine's last error (the thread's, without an engine)
// and returns `code`.
vt_status engine_error(vt_engine* e, vt_status code, const std::string& msg);

// Runs the body of a C entry point.  An exception -- std::bad_alloc from a
// builder's containers on a very large scene, std::length_error, anything a
// library throws -- must not unwind through extern "C" into ctypes or N-API:
// it becomes VT_ERR_OUT_OF_MEMORY (allocation) or VT_ERR_BUILD (any other
// exception), with its message.
template <class F>
vt_status vt_guarded(vt_engine* e, F&& body) noexcept {
  try {
    return body();
  } catch (const std::bad_alloc&) {
    try {
      return engine_error(e, VT_ERR_OUT_OF_MEMORY, "out of host memory");
    } catch (...) {
      return VT_ERR_OUT_OF_MEMORY;
    }
  } catch (const std::exception& ex) {
    try {
      return engine_error(e, VT_ERR_BUILD, std::string("internal error: ") + ex.what());
    } catch (...) {
      return VT_ERR_BUILD;
    }
  } catch (...) {
    return VT_ERR_BUILD;
  }
}

struct DeviceBuffer {
  void* ptr = nullptr;
  size_t bytes = 0;
  // Grows the allocation to at least `want` bytes (contents are not kept).
  vt_status ensure(vt_engine* e, size_t want);
  void release();
};

struct SceneBuffers {
  DeviceBuffer blas, tlas, tris, instances;          // traversal layout
  DeviceBuffer blas4, tlas4;                         // Node128 trees
  DeviceBuffer positions, normals, indices, tri_material, materials;  // shading data
  void release();
};

struct RenderState;      // integrator state (render.hip)
struct RenderSceneData;  // one scene's shading data (render.cpp)

// Test hooks (vt_debug_inject_failure): each bit makes the next matching step
// fail once -- packing throws std::bad_alloc, the shading-data upload reports
// an allocation failure, or its preparation throws std::runtime_error.
enum : uint32_t { kInjectPackThrow = 1u, kInjectRenderAlloc = 2u, kInjectRenderThrow = 4u };
struct CwbvhState;   // wide-BVH traversal buffers (cwbvh.cpp)
struct RefitState;   // device refit buffers (refit.cpp)
struct GpuBuildScratch;  // device BLAS builder buffers (gpubuild.cpp)
struct NrcState;     // neural radiance cache (nrc.hip)

struct vt_engine {
  int device = 0;
  vt_options opts{};
  hipStream_t stream = nullptr;
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  bool trace_timed = false;
  bool frame_timed = false;
  uint32_t cu_count = 0;
  uint32_t trace_blocks_per_cu = 1;
  uint32_t trace_grid = 1;
  uint32_t refill_min = 8;
  uint32_t trace_steps = 16;
  uint32_t near_first = 2;
  uint32_t shade_sort = 0;  // material-sorted shading order (vt_set_shade_sort)
  uint32_t defer_min = 12;
  uint32_t node_width = 4;  // 4: Node128 trees where the stack bound allows, 2: Node64
  uint32_t hit_filter = 0;  // vt_trace_*: 0 raw traversal, 1 the path tracer's sided wrappers
  bool counting = false;
  DeviceBuffer counter, error_word, spill;
  DeviceBuffer tmp_rays, tmp_out, tmp_counts;
  std::vector<void*> user_allocs;
  SceneBuffers scene;
  std::unique_ptr<vt::Pack> pack;
  std::unique_ptr<vt::DeviceLayout> layout;
  std::vector<vt_material> materials;
  float scene_radius = 0;
  bool has_scene = false;
  float last_frame_ms = 0;
  uint32_t samples_accumulated = 0;
  uint32_t last_ray_count = 0;
  RenderState* render = nullptr;
  NrcState* nrc = nullptr;
  CwbvhState* cwbvh = nullptr;
  std::string last_error;
  uint32_t fail_inject = 0;    // kInject* bits
  uint32_t build_threads = 0;  // host threads of the scene builder (0: every hardware thread)
  uint32_t device_build_min = 0;  // meshes with this many triangles build on the device (0: none)
  void* noalloc_scratch = nullptr;  // this device's copy of noalloc.hip's scratch array
  RefitState* refit = nullptr;
  GpuBuildScratch* gpubuild = nullptr;  // kept between builds, freed with the engine
  // Counts scene changes (vt_set_scene, refits): recorded frame graphs
  // depend on the scene they were recorded with.
  uint64_t scene_generation = 0;
};

// Looks up the device address of noalloc.hip's scratch array for the engine's
// device (noalloc.hip; called once by vt_create).
vt_status engine_noalloc_setup(vt_engine* e);

// Threads the scene builder uses for this engine.
uint32_t engine_build_threads(const vt_engine* e);
// pack_scene with the engine's builder settings: host threads, and the
// device builder for meshes of at least device_build_min triangles.
vt::BuildError engine_pack_scene(vt_engine* e, const vt::PrimitiveIn* prims, uint32_t count, vt::Pack& out);

vt_status engine_upload(vt_engine* e, DeviceBuffer& dst, const void* src, size_t bytes);
// Uploads the traversal buffers of a device layout (Node64 + Node128 trees,
// triangles, instances).
vt_status engine_upload_trace(vt_engine* e, const vt::DeviceLayout& layout, SceneBuffers& out);
// Device view of one scene's traversal buffers; `wide` reports whether its
// Node128 trees can be walked (their worst-case stack fits the kernel's).
vtk::DeviceScene engine_device_scene(const SceneBuffers& b, const vt::DeviceLayout& layout, bool* wide);
// Fills the scene and scheduling fields of a launch for the engine's scene.
vtk::TraceParams engine_trace_params(vt_engine* e);
// Packs, builds and uploads one scene's traversal buffers (engine.cpp).
vt_status engine_build_trace_scene(vt_engine* e, const vt_scene_desc* scene, SceneBuffers& out,
                                   vt::DeviceLayout& layout);
// The procedural sky's baked map (ProceduralSkyEnvironment -> Preetham bake,
// environmentPacking.ts:237-282): width x height rgb binary32 texels (render.cpp).
constexpr uint32_t kSkyWidth = 256, kSkyHeight = 128;
std::string engine_bake_sky(const vt_environment& env, uint32_t width, uint32_t height, std::vector<float>& rgb,
                            double* luminance_integral);
// Hooks into the integrator (render.cpp).  prepare builds and uploads a new
// scene's shading data without touching the bound scene; commit binds it
// (cannot fail; takes ownership); discard frees an unbound set.
vt_status engine_prepare_render(vt_engine* e, const vt::Pack& pack, const std::vector<vt_material>& materials,
                                const vt_environment& env, float scene_radius, const vt_emitter* emitters,
                                uint32_t emitter_count, RenderSceneData** out);
void engine_commit_render(vt_engine* e, RenderSceneData* d);
void engine_discard_render(RenderSceneData* d);
void engine_release_render(vt_engine* e);
// Recomputes the bound scene's ray bias and normalisation box from the
// pack's world bounds after a refit moved them (render.cpp).
void engine_render_rebound(vt_engine* e);
// Refit state (refit.cpp): per-scene device copies are dropped when the
// scene changes; everything is freed with the engine.
void engine_refit_scene_changed(vt_engine* e);
void engine_release_refit(vt_engine* e);
void engine_release_gpubuild(vt_engine* e);
void engine_release_nrc(vt_engine* e);
// Wide-BVH (cwbvh-closest) traversal state (cwbvh.cpp): built on first use
// per scene, released on scene change.
vt_status engine_cwbvh_ensure(vt_engine* e);
void engine_release_cwbvh(vt_engine* e);
// Device view of the wide forest (null before engine_cwbvh_ensure succeeds).
namespace vtw {
struct Scene;
}
const vtw::Scene* engine_cwbvh_view(vt_engine* e);
// Closest hit in the cwbvh-closest mode: 2 x float4 per ray in, 5 words per
// ray out, optional world hit positions; the ray count is read from d_n when
// it is set (n is then the capacity); sided = 1 applies traceClosest's
// single-sided re-trace.
vt_status engine_trace_cwbvh(vt_engine* e, const float4* d_rays, uint32_t n, const uint32_t* d_n,
                             uint32_t* d_hits, float* d_hit_pos, uint32_t sided);
// Cache hooks for the integrator (nrc.cpp): counts are device words.
vt_status engine_nrc_ensure(vt_engine* e);
// Creates the cache if it does not exist, without joining a training step.
vt_status engine_nrc_create(vt_engine* e);
// The path tracer's cache training step runs on a stream of its own: join
// makes the engine stream wait for it, quiesce waits for it on the host
// (before buffers it reads are freed).
vt_status engine_nrc_join(vt_engine* e);
vt_status engine_nrc_quiesce(vt_engine* e);
vt_status engine_nrc_infer_counted(vt_engine* e, const float* d_feat, const uint32_t* d_count, float* d_rgb,
                                   uint32_t capacity);
vt_status engine_nrc_train_counted(vt_engine* e, const float* d_feat, const float* d_tgt, const uint32_t* d_count,
                                   uint32_t capacity);
